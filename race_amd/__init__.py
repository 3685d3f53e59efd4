"""race_amd -- MI355X-native RACE execution engine.

The product is librace_amd.so (C ABI in include/race_amd.h, HIP kernels for
gfx950 under race_amd/csrc).  This package holds the Python mirror of
RACE::Interface over that ABI, the synthetic matrix generators and the build
script.
"""
from .interface import (BF16, FP32, FP64, K_GS, K_KACZ, K_MPK, K_MPK_HIGHEST, K_SPMTV, K_SPMV, K_SYMMSPMV,
                        LB_NNZ, LB_ROW, ONE, POWER, TWO, TWO_BLOCK, Batch, DeviceArray, Interface, RaceError,
                        device_count, lb_split, mpk_consolidate, mpk_hopeless_regions, rcm, split_threads, schedule_info, simdify_resolve, simdifyD1)

__all__ = ["Interface", "Batch", "DeviceArray", "RaceError", "lb_split", "split_threads", "mpk_hopeless_regions", "mpk_consolidate", "device_count", "simdify_resolve",
           "simdifyD1", "schedule_info", "rcm", "ONE", "TWO", "POWER", "K_SYMMSPMV", "K_GS", "K_KACZ", "K_SPMV", "K_MPK",
           "K_SPMTV", "K_MPK_HIGHEST", "FP64", "FP32", "BF16", "LB_NNZ", "LB_ROW", "TWO_BLOCK"]
