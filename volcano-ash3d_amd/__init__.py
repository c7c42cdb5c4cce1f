"""MI355X-native transport step of the Ash3d volcanic-ash model.

Host side of the drop-in for the reference's hot path (Set_BC, AdvectHorz /
advect_x / advect_y, AdvectVert / advect_z, DiffuseHorz / DiffuseVert with
diff_* and diffCN_*).  The compute runs in the HIP library built from csrc/
(libash3d_tr.so, C ABI declared in include/ash3d_tr.h); this package holds
the ctypes mirror of that ABI (``transport``), the host-side state in the
reference's array layouts (``state``) and the synthetic workloads
(``workloads``).
"""
from . import state, workloads  # noqa: F401
from .state import (LIM_BW, LIM_FROMM, LIM_LAXWEN, LIM_MC, LIM_MINMOD, LIM_NONE,  # noqa: F401
                    LIM_SUPERBEE, TransportState)
