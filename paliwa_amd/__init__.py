"""paliwa_amd -- MI355X (gfx950) lifting-wavelet transform with the interface of
freifrauvonbleifrei/paliwa's hierarchize / dehierarchize.

The compute path is the HIP library paliwa_amd/lib/libpaliwa_hip.so (C ABI:
include/paliwa_hip.h); this package is its Python host mirror.  Importing the
package fails if the library has not been built.
"""
from .transform import (  # noqa: F401
    Affine,
    FilterOffsetError,
    HipError,
    LevelError,
    PaliwaError,
    ShapeError,
    UnknownWaveletError,
    UnsupportedError,
    auto_path,
    dehierarchize,
    dehierarchize_checked,
    dehierarchize_in,
    dehierarchize_in_checked,
    hierarchize,
    hierarchize_checked,
    hierarchize_in,
    hierarchize_in_checked,
    lifting_wavelet_filter_offsets_and_coefficients,
    lifting_wavelet_reconstruct_offsets_and_coefficients,
    transform_in,
    transform_in_checked,
    wavelet_filters,
)
from .batch import CombinationBatch, combination_levels  # noqa: F401
from .combine import SparseGridCombination  # noqa: F401
from ._lib import LIBRARY_PATH  # noqa: F401

__version__ = "0.1.0"
