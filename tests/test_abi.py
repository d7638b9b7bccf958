"""CPU checks of the C-ABI boundary: the library loads without a GPU, exports every function that
include/rtc8x.h declares, the Python bindings declare exactly those functions, and the ctypes mirrors
of the ABI structs have the library's sizes."""
import ctypes as C
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "rtc8x.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    names = set()
    for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s+\**\s*(rtc8x_[a-z0-9_]+)\s*\(", text, flags=re.M):
        names.add(m.group(1))
    return names


def test_header_declares_the_boundary():
    names = declared_functions()
    for required in ("rtc8x_trace", "rtc8x_nrc_create", "rtc8x_nrc_destroy", "rtc8x_nrc_infer", "rtc8x_nrc_train",
                     "rtc8x_nrc_prepare", "rtc8x_nrc_propagate", "rtc8x_nrc_shuffle", "rtc8x_nrc_accumulate",
                     "rtc8x_sky_bake", "rtc8x_env_build_sampler", "rtc8x_grid_upload"):
        assert required in names, required
    assert len(names) >= 40


def test_library_exports_every_declared_symbol():
    from rtcamp8_amd._lib import LIB_PATH

    if not LIB_PATH.exists():
        pytest.skip("library not built")
    handle = C.CDLL(str(LIB_PATH))
    missing = [n for n in sorted(declared_functions()) if not hasattr(handle, n)]
    assert not missing, missing


def test_bindings_match_header():
    from rtcamp8_amd import _abi

    assert set(_abi.PROTOTYPES) == declared_functions()


def test_struct_layouts_and_version():
    from rtcamp8_amd import _abi
    from rtcamp8_amd._lib import LIB_PATH, lib

    if not LIB_PATH.exists():
        pytest.skip("library not built")
    h = lib()  # checks every struct size against rtc8x_sizeof and the ABI version
    assert h.rtc8x_abi_version() == 1
    for name, cls in _abi.STRUCTS.items():
        assert h.rtc8x_sizeof(name.encode()) == C.sizeof(cls), name
    assert h.rtc8x_sizeof(b"no_such_type") == 0


def test_errors_are_reported_not_thrown():
    """Argument validation runs before any device work, so it is testable without a GPU."""
    from rtcamp8_amd import _abi
    from rtcamp8_amd._lib import LIB_PATH, Rtc8xError, call, lib

    if not LIB_PATH.exists():
        pytest.skip("library not built")
    with pytest.raises(Rtc8xError) as e:
        call("rtc8x_trace", None, None, _abi.ENTRY_PATH_TRACE, 1)
    assert e.value.status != _abi.RTC8X_OK
    assert lib().rtc8x_last_error()
    with pytest.raises(Rtc8xError):
        call("rtc8x_nrc_create", None, None)
