"""The C ABI without a GPU: the library loads, exports every entry point that
include/lbm.h declares, rejects bad configurations with LBM_EINVAL before it
touches a device, and its host-side tables (inlet perturbation) are
bit-identical to the oracle's."""
import os
import re

import numpy as np
import pytest

import lbm_multigpu_amd as L
import oracle_lib as O
from lbm_multigpu_amd import config as cfgmod

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "lbm.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(lbm_[a-z0-9_]+)\s*\(", txt)))


def test_missing_library_fails_loudly():
    """With no built library the binding raises; nothing falls back to a CPU path."""
    saved = L._lib
    L._lib = None
    try:
        with pytest.raises(OSError, match="not built"):
            L.load_library(os.path.join(ROOT, "no_such_dir", "liblbm.so"))
    finally:
        L._lib = saved


def test_closure_setters_reject_a_null_context():
    lib = L.load_library()
    assert lib.lbm_set_closure_target(None, 1) == -22
    assert lib.lbm_set_closure_base(None, 1) == -22


def test_library_exports_every_declared_symbol():
    lib = L.load_library()
    names = header_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert set(L.EXPORTS) <= set(names)


def test_config_default_values():
    c = L.default_config()
    assert (c.nx, c.ny, c.nz) == (64, 64, 64)
    assert c.model == L.MODEL_BGK and c.storage == L.STORE_F32
    assert abs(c.s_dhum[1] - 1.19) < 1e-6 and abs(c.s_dhum[16] - 1.98) < 1e-6


@pytest.mark.parametrize("field,value,msg", [
    ("omega", 2.5, "omega"), ("omega", 0.0, "omega"), ("nx", 1, "domain"), ("model", 7, "model"),
    ("storage", 9, "storage"), ("cs", -1.0, "CS"), ("geometry", 5, "geometry"),
])
def test_create_rejects_bad_config(field, value, msg):
    c = L.default_config()
    setattr(c, field, value)
    with pytest.raises(L.LBMError) as e:
        L.Solver(c)
    assert e.value.code == -22 and msg in str(e.value)


def test_create_rejects_planes_beyond_32bit_offsets():
    """Kernels address a direction plane with 32-bit byte offsets, so the
    limit follows the padded plane: 2 x 2^14 x 2^14 is 2^29 nodes (within the
    node limit) but 2^34 padded ones (x padded from 2 to 64), over 4 GiB per
    plane in FP64 and in FP16 alike."""
    c = L.default_config()
    c.nx, c.ny, c.nz = 2, 1 << 14, 1 << 14
    for st in (L.STORE_F64, L.STORE_F16):
        c.storage = st
        with pytest.raises(L.LBMError) as e:
            L.Solver(c)
        assert e.value.code == -22 and "4 GiB" in str(e.value)


def test_cpu3d_model_rejects_les():
    c = L.default_config()
    c.model = L.MODEL_MRT_CPU3D
    c.cs = 0.01
    with pytest.raises(L.LBMError) as e:
        L.Solver(c)
    assert e.value.code == -22


def test_create_without_gpu_reports_device_error():
    try:
        import torch  # noqa: F401
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if has_gpu:
        pytest.skip("a GPU is present")
    with pytest.raises(L.LBMError) as e:
        L.Solver(L.default_config())
    assert e.value.code == -5


def test_inlet_profile_bit_identical_to_oracle():
    for seed in (0, 1, 12345):
        a = L.inlet_profile(128, 128, 0.05, 1e-3, seed)
        b = O.inlet_profile(128, 128, 0.05, 1e-3, seed)
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    lib = L.load_library()
    for c in (0, 1, 2, 1000, 2 ** 40):
        assert lib.lbm_hash_u32(7, c) == O.lib().orc_hash_u32(7, c)


def test_inputs_round_trip_and_reference_record():
    """A run record of the reference (MRTtest/mrt_les.inputs, as written by
    WriteInputs) parses to the same case, and our writer round-trips."""
    rec = ("Base domain size \t64x124x62\nBase blocksize: \t64x1x1\nObst1 location: \t(31.5,92.5,32.5)\n"
           "Obst1 radius: \t31\nObst2 location: \t(25.5,25.5,32.5)\nObst2 radius: \t10\nRE: \t5400\n"
           "UMAX: \t0.06\nomega \t: 1.99177\nDPDY \t: -5.16e-07\nTMAX: \t1000\nSTARTF: \t3000000\n"
           "START_VELAV: \t200000\nSTART_VELFLUC: \t1600000\nREFINEMENT: \t0\nMODEL: \tMRT\n"
           "Smagorinsky LES: \t1\nCS: \t0.02\nLR domain size \t128x80x8\nLR factor \t0.5\n"
           "LR location \t128.25x64.25x-0.75\nLR blocksize: \t64x1x1\nomega in LR \t: 1.9836\n"
           "GPUs per node \t: 1\n")
    c = cfgmod.read_inputs(rec)
    assert (c.nx, c.ny, c.nz) == (64, 124, 62)
    assert c.obst1 == (31.5, 92.5, 32.5) and c.obst1_r == 31
    assert c.re == 5400 and abs(c.umax - 0.06) < 1e-12 and abs(c.omega - 1.99177) < 1e-12
    assert abs(c.dpdy + 5.16e-7) < 1e-15 and c.tmax == 1000 and c.startf == 3000000
    assert c.model == "MRT" and c.les and abs(c.cs - 0.02) < 1e-12
    assert c.lr_size == (128, 80, 8) and c.lr_loc == (128.25, 64.25, -0.75) and c.omega_lr == 1.9836
    # omega derived from RE when the record has none (MRTtest/DynInlet3.cu:2869)
    c2 = cfgmod.read_inputs(rec.replace("omega \t: 1.99177\n", ""))
    assert abs(c2.omega - 1.99177) < 1e-5
    assert abs(cfgmod.omega_lr(c2.omega) - 1.9836) < 1e-4
    # round trip of every field we write
    ch = cfgmod.channel(force_steps=10)
    back = cfgmod.read_inputs(cfgmod.write_inputs(ch))
    for f in ("nx", "ny", "nz", "bc", "model", "mrt_rates", "les", "omega", "umax", "halfway",
              "storage", "geometry", "obst1", "obst1_r", "inlet_amp", "seed", "force_steps"):
        a, b = getattr(ch, f), getattr(back, f)
        if isinstance(a, float):
            assert abs(a - b) <= 1e-5 * max(1, abs(a)), f
        elif isinstance(a, tuple) and not isinstance(a[0], str):
            assert all(abs(float(x) - float(y)) < 1e-5 for x, y in zip(a, b)), f
        else:
            assert a == b, f


def test_inputs_parses_reference_separators():
    """The three separator styles of the reference ("Key: \\tv", "Key \\t: v",
    "Key \\tv") and a truncated record."""
    c = cfgmod.read_inputs("Base domain size \t10x20x30\nUMAX: \t0.1\nomega \t: 1.5\n")
    assert (c.nx, c.ny, c.nz) == (10, 20, 30) and c.umax == 0.1 and c.omega == 1.5
    c = cfgmod.read_inputs("Base domain size \t10x20x30\nUMAX:")
    assert c.nx == 10


def _kernel_private_segments():
    """(kernel name, private segment bytes per lane) of every kernel in the
    built library's gfx950 code object (offload bundle -> AMDGPU metadata)."""
    import re
    import struct
    import subprocess
    data = open(L.LIB_PATH, "rb").read()
    i = data.find(b"__CLANG_OFFLOAD_BUNDLE__")
    assert i >= 0, "no offload bundle in liblbm.so"
    n = struct.unpack_from("<Q", data, i + 24)[0]
    p, code = i + 32, None
    for _ in range(n):
        off, size, tl = struct.unpack_from("<QQQ", data, p)
        triple = data[p + 24:p + 24 + tl].decode()
        p += 24 + tl
        if triple.endswith("gfx950"):
            code = data[i + off:i + off + size]
    assert code, "no gfx950 code object"
    import tempfile
    with tempfile.NamedTemporaryFile(suffix=".co") as fh:
        fh.write(code)
        fh.flush()
        out = subprocess.run(["/opt/rocm/llvm/bin/llvm-readelf", "--notes", fh.name], capture_output=True,
                             text=True, check=True).stdout
    names = re.findall(r"\.name:\s+(\S+)", out)
    priv = [int(v) for v in re.findall(r"\.private_segment_fixed_size:\s+(\d+)", out)]
    assert len(names) == len(priv) > 50
    return list(zip(names, priv))


def test_kernels_keep_their_state_in_registers():
    """No kernel of the library spills to scratch (a spilled population array
    cost the stream-collide a quarter of its bandwidth once)."""
    spills = [(n, b) for n, b in _kernel_private_segments() if b]
    assert not spills, spills
