# Portions of this file reproduce EASIER (jack-pan-ai/EASIER),
# Copyright (c) Microsoft Corporation, MIT License; the licence text is in
# easier_amd/THIRD_PARTY_NOTICES.txt.
"""Linear-system operator (easier/numeric/linsys.py:10-24)."""
from torch import nn


class Linsys(nn.Module):
    """y = reducer(selector(x) * Ae) + x * Av: an SpMV whose off-diagonal
    entries are the edge values Ae (row = reducer idx, column = selector idx)
    and whose diagonal is Av."""

    def __init__(self, Av, Ae, selector, reducer) -> None:
        super().__init__()
        self.Av = Av
        self.Ae = Ae
        self.selector = selector
        self.reducer = reducer

    def forward(self, x):
        return self.reducer(self.selector(x) * self.Ae) + x * self.Av
