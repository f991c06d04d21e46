"""mcnla_amd: MI355X-native integrated randomized SVD (iSVD) behind MCNLA's stage-plugin API.

The compute lives in libmcnla_amd.so (hand-written HIP kernels for gfx950 behind the C ABI in
include/mcnla_amd/isvd_capi.h).  This package is the Python host mirror of the reference's
``mcnla::isvd`` classes; see :mod:`mcnla_amd.isvd`.
"""
from . import _lib
from ._lib import McnlaError, lib
from .device import DeviceArray, device_count, mem_info, synchronize
from .isvd import (
    ColBlockGaussianProjectionSketcher, ColBlockGramianFormer, CollectionFromPartialSumToRowBlockConverter,
    ColumnSamplingSketcher, RowBlockColumnSamplingSketcher, column_indices_reference,
    CollectionFromRowBlockConverter, CollectionToRowBlockConverter, DummyConverter,
    GaussianProjectionSketcher, GramianFormer, GramianOrthogonalizer, KolmogorovNagumoIntegrator,
    MatrixFromColBlockToAllConverter, MatrixFromRowBlockConverter, MatrixFromRowBlockToAllConverter,
    MatrixToRowBlockConverter, Parameters, RowBlockGaussianProjectionSketcher,
    RowBlockExtrinsicMeanIntegrator, RowBlockGramianFormer, RowBlockGramianKolmogorovNagumoIntegrator,
    RowBlockGramianOrthogonalizer, RowBlockGramianWenYinIntegrator,
    RowBlockKolmogorovNagumoIntegrator, RowBlockReductionIntegrator, RowBlockSymmetricFormer,
    RowBlockWenYinIntegrator, Solver, SvdFormer, gaussian, solve_batch, synth_lowrank, tune_clear,
    tune_get, tune_load, tune_save, tune_set, tuning_default,
)
from .serve import Served, Server, required_bytes
from . import workloads

PREC_F64 = _lib.PREC_F64
PREC_MIXED_BF16 = _lib.PREC_MIXED_BF16
PREC_MIXED_BF16X3 = _lib.PREC_MIXED_BF16X3
PREC_MIXED_F16 = _lib.PREC_MIXED_F16
PREC_MIXED_F16X2 = _lib.PREC_MIXED_F16X2
PREC_MIXED_F16X3 = _lib.PREC_MIXED_F16X3
OMEGA_PHILOX = _lib.OMEGA_PHILOX
OMEGA_REFERENCE = _lib.OMEGA_REFERENCE
OMEGA_PHILOX_F16 = _lib.OMEGA_PHILOX_F16
ROUTE_AUTO = _lib.ROUTE_AUTO
ROUTE_SKETCH = _lib.ROUTE_SKETCH
ROUTE_GRAM_OF_A = _lib.ROUTE_GRAM_OF_A

__version__ = "0.1.0"
