"""The built HIP library loads without a GPU and exports every entry point the
C headers under include/ declare (no device calls are made here)."""
import ctypes
import glob
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "inspire-rs_amd", "libinspire_gpu.so")


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        for m in re.finditer(r"^\s*(?:int|void|size_t|uint32_t|uint64_t)\s+\**\s*(inspire_\w+)\s*\(", text, re.M):
            names.add(m.group(1))
    return names


def test_header_declares_the_api():
    names = declared_functions()
    for must in ("inspire_ctx_create", "inspire_shard_precompute_packing", "inspire_seat_create",
                 "inspire_respond_batch", "inspire_respond_batch_device"):
        assert must in names


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "build the HIP library first (__graft_entry__.build())"
    lib = ctypes.CDLL(LIB)
    missing = [n for n in sorted(declared_functions()) if not hasattr(lib, n)]
    assert not missing, missing
