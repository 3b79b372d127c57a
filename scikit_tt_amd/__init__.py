"""scikit_tt_amd: MI355X-native TT-ALS (drop-in for scikit_tt's ALS path).

    from scikit_tt_amd.tensor_train import TT
    import scikit_tt_amd.solvers.sle as sle
    import scikit_tt_amd.solvers.evp as evp

Every compute call goes to libttals.so (hand-written HIP for gfx950).
"""
from .tensor_train import TT
from . import solvers

__version__ = '0.1.0'
