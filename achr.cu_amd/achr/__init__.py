"""MI355X ACHR sampler: Python view of libachr.so (include/achr.h)."""
from ._lib import CLI_PATH, LIB_PATH, WARMUP_CLI_PATH, AchrError, lib
from .sampler import (PROJ_FP64, PROJ_NULL_BF16X3, PROJ_NULL_F32, PROJ_RES_BF16X3, PROJ_RES_F32, Batch, DirNet,
                      Model, Options, Sampler, bank_reciprocal, gemm_f32, gemm_sub_f64, philox4x32_10, read_points_bin,
                      lp_polish, quotient, read_warmup_csv, warmup_points, write_points_bin, write_points_csv)

__all__ = ["AchrError", "Batch", "CLI_PATH", "WARMUP_CLI_PATH", "DirNet", "PROJ_FP64", "bank_reciprocal", "PROJ_NULL_BF16X3", "PROJ_NULL_F32", "PROJ_RES_BF16X3",
           "PROJ_RES_F32", "gemm_f32", "gemm_sub_f64", "LIB_PATH", "Model", "Options", "Sampler", "lib", "philox4x32_10", "lp_polish", "quotient",
           "read_points_bin", "read_warmup_csv", "warmup_points", "write_points_bin", "write_points_csv"]
