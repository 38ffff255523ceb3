"""MI355X Openpipeflow time-step path: Python host mirror of the reference
interface over the HIP library libopf.so (include/opf.h)."""
from .engine import Context, bytes_needed, host_mesh, LDP, LDM, LDZ, LNP, DIAG_FIELDS  # noqa: F401
from .engine import state_file_info, read_state_file, write_state_file  # noqa: F401
from ._lib import OpfError  # noqa: F401
from . import ic  # noqa: F401
from .serve import Server  # noqa: F401
