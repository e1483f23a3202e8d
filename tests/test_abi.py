"""The C ABI without a device: the library loads, exports every function the
public headers declare, and refuses bad arguments before touching a device."""
import ctypes
import glob
import os
import re

import numpy as np

from conftest import ROOT


def declared_functions(headers=None):
    names = set()
    paths = ([os.path.join(ROOT, "include", h) for h in headers] if headers else
             [h for h in glob.glob(os.path.join(ROOT, "include", "*.h"))
              if not h.endswith("spotfinder_nxmx.h")])
    for h in paths:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(sf_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_headers_declare_the_surface():
    names = declared_functions()
    for must in ("sf_find_run", "sf_find_fetch", "sf_available", "sf_backend",
                 "sf_ws_threshold", "sf_synth_frame"):
        assert must in names


def test_library_exports_every_declared_function(sf):
    lib = ctypes.CDLL(sf.LIBRARY)
    missing = [n for n in sorted(declared_functions()) if not hasattr(lib, n)]
    assert not missing, missing


def test_reader_library_exports_its_header(sf):
    """libspotfinder_nxmx.so (built where HDF5 is) exports every function
    include/spotfinder_nxmx.h declares."""
    import pytest
    if not os.path.exists(sf.NXMX_LIBRARY):
        pytest.skip("the NXmx reader is not built (no HDF5 headers)")
    lib = ctypes.CDLL(sf.NXMX_LIBRARY)
    names = declared_functions(["spotfinder_nxmx.h"])
    assert "sf_nxmx_read_chunks" in names
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_backend_and_reason(sf):
    assert sf.backend() == "HIP"
    if not sf.available():
        assert sf.unavailable_reason() != ""
    else:
        assert sf.unavailable_reason() == ""


def test_bad_arguments_are_refused_without_a_device(sf):
    L = sf.lib()
    found = ctypes.c_size_t(7)
    frame = np.zeros((4, 4), np.uint32)
    p = frame.ctypes.data_as(ctypes.c_void_p)
    assert L.sf_find_run(32, None, 4, 4, ctypes.byref(found)) == -1
    assert found.value == 0
    assert L.sf_find_run(32, p, 0, 4, ctypes.byref(found)) == -1
    assert L.sf_find_run(32, p, 4, 0, ctypes.byref(found)) == -1
    assert L.sf_find_run(32, p, 1 << 20, 1 << 12, ctypes.byref(found)) == -1
    assert L.sf_find_run(8, p, 4, 4, ctypes.byref(found)) == -1
    assert sf.find(np.zeros((0, 5), np.uint16))[0] == -1


def test_record_layouts_match_the_reference():
    from oracle import oracle as O
    import spotfinder_hip as sfm
    assert sfm.SIGNAL_PIXEL == O.SIGNAL_PIXEL and sfm.SIGNAL_PIXEL.itemsize == 16
    assert sfm.PIXEL.itemsize == 12 and sfm.SPOT.itemsize == 104
    ref = O.reference()
    if ref is not None:
        assert ref.lib.ref_sizeof_spot() == 104


def test_synthetic_frames_are_deterministic(sf):
    d = sf.synth_desc((300, 200), seed=5, spots_per_frame=20)
    a = sf.synth_frame(d, 3)
    d.threads = 1
    b = sf.synth_frame(d, 3)
    assert a.tobytes() == b.tobytes()
    assert sf.synth_frame(d, 4).tobytes() != a.tobytes()
    g = sf.synth_frame(sf.synth_desc("4m"), 0)
    assert g.shape == (2162, 2068)
    # module gaps: 12 columns after each 1028, 38 rows after each 512
    assert (g[:, 1028:1040] == 0xFFFFFFFF).all()
    assert (g[512:550, :] == 0xFFFFFFFF).all()
    assert ((g == 65535).sum() > 100)


def test_cpp_layer_exports_the_gpu_interface():
    """libspotfinder_gpu.so defines every gpu:: function the drop-in header
    declares: the backend part and the standalone policy part."""
    import subprocess
    so = os.path.join(ROOT, "dials-metal_amd", "spotfinder_hip",
                      "libspotfinder_gpu.so")
    out = subprocess.run(["nm", "-DC", so], capture_output=True, text=True,
                         check=True).stdout
    defined = set()
    for line in out.splitlines():
        m = re.match(r"^[0-9a-f]+ [TW] (?:int )?(gpu::[^(\[]+)", line)
        if m:
            defined.add(m.group(1))
    for name in ("backend", "name", "stage0_window", "stage2_window", "fused",
                 "profile_stages", "last_stage_times", "report_windows",
                 "available", "unavailable_reason", "memory_free", "host_alloc",
                 "host_free", "internal::set_stage_times",
                 "internal::announce_windows", "internal::announce_fused",
                 "internal::default_stage0_window",
                 "internal::default_stage2_window",
                 "find<unsigned short>", "find<unsigned int>"):
        assert "gpu::" + name in defined, name


def test_backend_builds_against_the_reference_headers():
    """The HIP backend compiles, unchanged, against the reference's own
    src/spots/dext_gpu.hh and dext_gpu_internal.hh -- the headers it would be
    built with under -DSPOTFINDER_HIP=ON (syntax only, nothing written)."""
    import shutil
    import subprocess
    import pytest
    ref = "/root/reference/src/spots"
    if not os.path.exists(os.path.join(ref, "dext_gpu.hh")) or not shutil.which("g++"):
        pytest.skip("the reference sources are not on this machine")
    src = os.path.join(ROOT, "dials-metal_amd", "csrc", "dext_gpu_hip.cc")
    r = subprocess.run(["g++", "-std=c++20", "-fsyntax-only", "-I" + ref,
                        "-I" + os.path.join(ROOT, "include"), src],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_policy_environment_setters_and_announcements():
    """gpu::'s backend-independent policy (tests/cpp/test_policy.cc), which
    needs no device: the SPOTFINDER_GPU_* words, the warning for any other
    value, setters over the environment, and an announcement only when the
    configuration changes while reporting is on (dext_gpu.hh:50-134)."""
    import subprocess
    exe = os.path.join(ROOT, "tests", "cpp", "test_policy")
    cases = [({}, ("tile", "tile", "0"), []),
             ({"SPOTFINDER_GPU_STAGE0": "direct", "SPOTFINDER_GPU_FUSED": "1"},
              ("direct", "tile", "1"), []),
             ({"SPOTFINDER_GPU_STAGE2": "direct", "SPOTFINDER_GPU_FUSED": "0"},
              ("tile", "direct", "0"), []),
             ({"SPOTFINDER_GPU_STAGE0": "Tile", "SPOTFINDER_GPU_FUSED": "yes"},
              ("tile", "tile", "0"),
              ["warning: SPOTFINDER_GPU_STAGE0=Tile is not 'tile' or 'direct'; "
               "leaving it alone",
               "warning: SPOTFINDER_GPU_FUSED=yes is not '1' or '0'; leaving it alone"])]
    for env, expect, warnings in cases:
        full = {k: v for k, v in os.environ.items() if not k.startswith("SPOTFINDER_GPU_")}
        full.update(env)
        r = subprocess.run([exe, *expect], env=full, capture_output=True, text=True)
        assert r.returncode == 0, (env, r.stdout)
        err = r.stderr.splitlines()
        start = err.index("<<announce")
        assert err[:start] == warnings, env
        assert err[start + 1:err.index("announce>>")] == [
            "HIP: the three stages fused, one kernel",
            "HIP: stage0 direct window, stage2 direct window",
            "HIP: the three stages fused, one kernel"], env


def test_host_copy_pool_under_thread_sanitizer():
    """The host copy pool that decoders and the device grouping's drain share,
    driven from four threads at once, built with ThreadSanitizer and run
    (`make -C dials-metal_amd check-tsan`): every copy exact, no race."""
    import subprocess
    import pytest
    if not os.path.exists("/opt/rocm/llvm/bin/clang++"):
        pytest.skip("no clang++ with ThreadSanitizer's runtime")
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "dials-metal_amd"),
                        "check-tsan"], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "four threads: ok" in r.stdout
