"""C-ABI boundary checks that need no GPU: the library loads, exports every
entry point include/efvm.h declares, and fails loudly without a device."""
import ctypes
import os
import re

import pytest

import dev_field_efvm_amd as efvm

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_entry_points():
    text = open(os.path.join(ROOT, "include", "efvm.h")).read()
    return sorted(set(re.findall(r"\b(efvm3?_[a-z_0-9]+)\s*\(", text)))


def test_header_declarations_match_binding_list():
    assert declared_entry_points() == sorted(efvm.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = efvm.load_library()
    for name in declared_entry_points():
        assert hasattr(lib, name), name
        assert ctypes.cast(getattr(lib, name), ctypes.c_void_p).value


def test_product_never_imports_the_oracle():
    pkg = os.path.join(ROOT, "dev_field_efvm_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".cpp", ".hip", ".h")):
                src = open(os.path.join(dirpath, f)).read()
                assert "oracle" not in src.replace("oracle restatement", ""), f


def test_context_without_device_raises():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    with pytest.raises(efvm.EfvmError) as e:
        efvm.Context(8, 8)
    assert e.value.status == 6   # EFVM_ERR_NO_DEVICE


def test_half_periodic_registry_rejected_on_host():
    reg = efvm.BoundaryRegistry().with_(efvm.BoundarySide.XLo, efvm.Periodic())
    with pytest.raises(ValueError):
        efvm.StructuredBoundaryPlugin(reg)


def test_learned_mode_setters_reject_bad_arguments_without_a_device():
    """The learned-mode switches check their arguments before touching a context: a null
    context is EFVM_ERR_ARG (1) for the setters and the getter, with no device needed."""
    import ctypes as C
    from dev_field_efvm_amd.efvm import _bind_mlp
    lib = efvm.load_library()
    _bind_mlp(lib)
    assert lib.efvm_set_learned_bypass(None, 1) == 1
    assert lib.efvm_set_learned_fused(None, 1) == 1
    a, f = C.c_int(-1), C.c_int(-1)
    assert lib.efvm_get_learned_fused(None, C.byref(a), C.byref(f)) == 1
    assert (a.value, f.value) == (-1, -1)
