"""pmarch: MI355X ray marcher for SPH particle fluids (Python side of libpmarch.so)."""
from ._lib import (LIB_PATH, MODE_CONFIG, MODE_REFERENCE, NORMAL_CDIFF, NORMAL_GRADW, SKIP_DG_BOX,  # noqa: F401
                   SKIP_DG_CONS, SKIP_EMPTY, SKIP_NONE, Marcher, PmError, bgeo_read, header_functions, load)
from . import camera, scenes  # noqa: F401
from .dataset import dataset_files, read_dataset  # noqa: F401
