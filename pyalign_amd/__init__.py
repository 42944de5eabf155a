"""pyalign_amd: pyalign's pairwise DP alignment solver on the AMD MI355X.

The public names match pyalign (gaps, problems, solve, simple); the DP
kernels run on gfx950 through libpyalign_hip.so.
"""

__version__ = "0.1.0"

from .gaps import *  # noqa: F401,F403
from .solve import *  # noqa: F401,F403
from .problems import *  # noqa: F401,F403
from .simple import global_alignment, semiglobal_alignment, local_alignment  # noqa: F401
