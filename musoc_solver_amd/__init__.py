"""MI355X-native SEVP barotropic-pressure solver (TIMCOM's REPBIR / REP1 /
PRE / MATINV path), HIP kernels behind a C ABI (include/sevp.h)."""
from . import synth  # noqa: F401
from .sevp import (FORM_BIR, FORM_FULL, FORM_FULL_PERIODIC, PREC_BF16, PREC_FP32,  # noqa: F401
                   PREC_FP64, SevpError, SevpPlan, lib, matinv, pre, pre1, rep, rep1, repbir,
                   repbir_period, version)

__all__ = ["synth", "SevpPlan", "SevpError", "lib", "repbir", "rep1", "rep", "repbir_period",
           "pre", "pre1", "matinv", "version", "FORM_BIR", "FORM_FULL", "FORM_FULL_PERIODIC",
           "PREC_FP64", "PREC_FP32", "PREC_BF16"]
