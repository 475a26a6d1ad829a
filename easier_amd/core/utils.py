"""Logging and the compile-time exception type.

Mirrors easier/core/utils.py:9-34: the logger is named "easier.init", its
level comes from EASIER_LOG_LEVEL, `init_logger(rank)` renames it "Rank{r}".
"""
import logging
import os

logger: logging.Logger = logging.getLogger("easier.init")
logger.setLevel(os.environ.get("EASIER_LOG_LEVEL", logging.INFO))
_handler = logging.StreamHandler()
_handler.setFormatter(logging.Formatter(
    "%(asctime)s - %(name)s - %(levelname)s - %(message)s"))
logger.addHandler(_handler)


def init_logger(rank: int):
    logger.name = f"Rank{rank}"


class EasierJitException(Exception):
    """Raised for malformed programs and bad compile() arguments."""
