"""Programs and inputs of the benchmark configurations, written against the
public esr API the way the reference tutorials are (tutorial/poisson,
tutorial/shallow_water_equation), with in-memory tensors in place of HDF5
files."""
