"""cosmosim — MI355X TreePM gravity (HIP kernels behind the C ABI of
include/chui_treepm.h). Import path: add the chui-dev_amd directory to sys.path."""
from . import _lib, params
from ._lib import ChuiError, Options, default_options, set_param
from .treepm import ForceAccumulator, TreePmCoordinator, kdk_factors, split_diagnostics

__all__ = ["ChuiError", "Options", "default_options", "set_param", "ForceAccumulator", "TreePmCoordinator",
           "kdk_factors", "split_diagnostics", "_lib", "params"]
