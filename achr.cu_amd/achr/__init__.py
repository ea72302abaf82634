"""MI355X ACHR sampler: Python view of libachr.so (include/achr.h)."""
from ._lib import CLI_PATH, LIB_PATH, AchrError, lib
from .sampler import Model, Options, Sampler, philox4x32_10, read_warmup_csv, write_points_csv

__all__ = ["AchrError", "CLI_PATH", "LIB_PATH", "Model", "Options", "Sampler", "lib", "philox4x32_10",
           "read_warmup_csv", "write_points_csv"]
