"""MI355X-native DPG local assembly (the per-cell work of Camellia's
TBF::localStiffnessMatrixAndRHS) behind a C ABI; see include/dpg_capi.h."""
from . import capi  # noqa: F401
