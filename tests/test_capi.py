"""The C ABI library: loads without a GPU, exports every entry point that
include/su2mc.h declares, and validates arguments on the host (no kernel is
launched by these tests)."""
import ctypes as C
import os
import re
import subprocess

import pytest

import qcd_2d_gpu_amd as q
from qcd_2d_gpu_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "su2mc.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(su2_[a-z_0-9]+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declarations_found():
    names = declared_functions()
    assert "su2_mc_up_per" in names and "su2_en_av" in names and "su2_run" in names
    assert len(names) >= 40


def test_library_exports_every_declared_symbol():
    lib = q.load_library()
    out = subprocess.run(["nm", "-D", "--defined-only", L.LIB_PATH], capture_output=True, text=True, check=True)
    exported = set(re.findall(r" T (su2_[a-z_0-9]+)", out.stdout))
    missing = [n for n in declared_functions() if n not in exported]
    assert not missing, missing
    for n in declared_functions():
        assert hasattr(lib, n)
    # the Python binding covers the whole header
    assert set(declared_functions()) == set(L.PROTOTYPES)


def test_struct_layout_matches_reference_struct():
    # struct lattice (functions/lattice_setup.h:2-14): 7 ints, 2 floats, 1 int
    assert C.sizeof(L.Su2Lattice) == 40
    assert [f[0] for f in L.Su2Lattice._fields_] == ["y", "x", "t", "V", "d", "dg", "dt", "beta", "epsilon",
                                                    "updates"]


def test_lattice_init_and_sizes():
    lib = q.load_library()
    l = q.make_lattice(128, 128, 128, 5.0)
    assert (l.t, l.x, l.y, l.V, l.d, l.dg, l.dt) == (128, 128, 128, 2097152, 3, 4, 12)
    assert lib.su2_field_bytes(C.byref(l), 0) == 12 * 2097152 * 8
    assert lib.su2_field_bytes(C.byref(l), 1) == 12 * 2097152 * 4
    assert lib.su2_rng_bytes(C.byref(l)) == 24 * 1048576
    assert lib.su2_accept_bytes(C.byref(l)) == 4 * 2097152
    assert lib.su2_en_av_work_bytes(C.byref(l)) == 16 * 4096  # two doubles per block of 256 half-sites


@pytest.mark.parametrize("ext", [(15, 16, 16), (16, 15, 16), (16, 16, 3), (0, 4, 4), (2, 2, -2)])
def test_odd_or_bad_extents_rejected(ext):
    with pytest.raises(q.Su2Error) as e:
        q.make_lattice(*ext, 3.0)
    assert e.value.code == L.SU2_EODD_EXTENT


def test_invalid_arguments_rejected_before_launch():
    lib = q.load_library()
    l = q.make_lattice(8, 8, 8, 3.0)
    lp = C.byref(l)
    fake = C.c_void_p(0x1000)
    assert lib.su2_mc_up_per(lp, None, 0, fake, fake, 1, 0, None) == L.SU2_EINVAL
    assert lib.su2_mc_up_per(lp, fake, 7, fake, fake, 1, 0, None) == L.SU2_EINVAL
    assert lib.su2_mc_up_per(lp, fake, 0, fake, fake, 5, 0, None) == L.SU2_EINVAL
    assert lib.su2_mc_up_per(lp, fake, 0, fake, None, 1, 0, None) == L.SU2_EINVAL
    assert lib.su2_mc_update(lp, fake, 0, fake, fake, 0, 1, 0, 3, 0, None) == L.SU2_EINVAL
    assert lib.su2_mc_update(lp, fake, 0, fake, fake, 3, 1, 0, 0, 0, None) == L.SU2_EINVAL
    assert lib.su2_en_av(lp, fake, 0, 0, None, fake, fake, None) == L.SU2_EINVAL
    bad = q.make_lattice(8, 8, 8, 3.0)
    bad.V = 7
    assert lib.su2_normalize(C.byref(bad), fake, 0, None) == L.SU2_EINVAL
    assert "V = 7" in lib.su2_last_error().decode()
    lu, bu = C.c_double(), C.c_double()
    for args in ((None, 8, 8, 10), (fake, 0, 8, 10), (fake, 2000, 8, 10), (fake, 8, 0, 10), (fake, 8, 8, 0)):
        assert lib.su2_phase_probe(*args, None, C.byref(lu), C.byref(bu), None) == L.SU2_EINVAL
    assert lib.su2_set_sweep_block(100) == L.SU2_EINVAL
    assert lib.su2_set_sweep_block(0) == L.SU2_OK
    assert lib.su2_strerror(L.SU2_EODD_EXTENT) == b"lattice extents must be even"


def test_couplings_that_would_never_finish_or_give_nan_are_rejected():
    """beta <= 0 would keep the Kennedy-Pendleton loop rejecting for ever and
    beta below 0.01 nearly so (about 1.26 (beta |A|)^1.5 accepted per trial),
    |eps| > 1 makes the Metropolis proposal NaN: refused on the host."""
    lib = q.load_library()
    fake = C.c_void_p(0x1000)
    for beta, eps, algo, ok in [(0.0, 0.1, 1, False), (-1.0, 0.1, 1, False), (float("nan"), 0.1, 1, False),
                                (0.0099, 0.1, 1, False), (1e-6, 0.1, 1, False), (0.01, 0.1, 1, True),
                                (float("inf"), 0.1, 0, False), (2.0, 1.5, 0, False), (2.0, -0.1, 0, False),
                                (0.0, 0.5, 0, True), (-1.0, 0.5, 0, True), (2.0, 1.0, 0, True)]:
        l = q.make_lattice(8, 8, 8, 1.0)
        l.beta, l.epsilon = beta, eps
        lp = C.byref(l)
        # no RNG pointer: nothing can be launched; valid couplings fail on the pointer instead
        assert lib.su2_mc_up_per(lp, fake, 0, fake, None, algo, 0, None) == L.SU2_EINVAL
        msg = lib.su2_last_error()
        assert lib.su2_mc_update(lp, fake, 0, fake, None, 0, algo, 0, 1, 0, None) == L.SU2_EINVAL
        assert lib.su2_last_error() == msg
        if ok:
            assert b"RNG states pointer is NULL" in msg, (beta, eps, algo, msg)
        else:
            assert b"beta" in msg or b"epsilon" in msg, (beta, eps, algo, msg)
    j = q.Job(8, 8, 8, 0.0)
    assert lib.su2_job_bytes(C.byref(j.c()), 0) == 0
    assert lib.su2_job_bytes(C.byref(q.Job(8, 8, 8, 0.0099).c()), 0) == 0
    assert lib.su2_job_bytes(C.byref(q.Job(8, 8, 8, 0.01).c()), 0) > 0
    j = q.Job(8, 8, 8, 2.0, algo=q.ALGO_METROPOLIS, epsilon=1.5)
    assert lib.su2_job_bytes(C.byref(j.c()), 0) == 0


def test_no_torch_in_product_package():
    """The product talks to HIP through libsu2mc only."""
    pkg = os.path.join(ROOT, "qcd_2d_gpu_amd")
    for fn in os.listdir(pkg):
        if fn.endswith(".py"):
            src = open(os.path.join(pkg, fn)).read()
            assert "import torch" not in src
            assert "oracle" not in re.sub(r"#.*", "", src).lower().replace("oracle/", ""), fn


def test_handle_api_rejects_cpu_backend_and_null_handles():
    """No CPU fallback: the CPU backend is refused before any device call."""
    lib = q.load_library()
    c = C.c_void_p(123)
    assert lib.su2_ctx_create(L.BACKEND_CPU, 0, C.byref(c)) == L.SU2_EINVAL
    assert c.value is None
    assert b"no CPU backend" in lib.su2_last_error()
    assert lib.su2_ctx_create(5, 0, C.byref(c)) == L.SU2_EINVAL
    h = C.c_void_p()
    assert lib.su2_lattice_create(None, 4, 4, 4, 0, 0, C.byref(h)) == L.SU2_EINVAL
    for name in ["su2_renormalize", "su2_accept_reset"]:
        assert getattr(lib, name)(None) == L.SU2_EINVAL
    assert lib.su2_rng_init(None, 1) == L.SU2_EINVAL
    assert lib.su2_sweep(None, 1, 0, 1.0, 0.1, 1) == L.SU2_EINVAL
    assert lib.su2_lattice_destroy(None) == 0
    assert lib.su2_ctx_destroy(None) == 0


def test_size_limit():
    """3V and every padded launch must stay below 2^32: the largest lattice
    has SU2_MAX_LATTICE_SITES = 1 431 633 920 sites."""
    lib = q.load_library()
    l = L.Su2Lattice()
    # 1 431 633 920 = 2 * 2 * 357 908 480 is accepted, 8 more sites are not
    assert lib.su2_lattice_init(C.byref(l), 2, 2, 357908480, 1.0, 0.1, 1) == 0
    assert l.V == 1431633920
    assert lib.su2_lattice_init(C.byref(l), 2, 2, 357908482, 1.0, 0.1, 1) == L.SU2_EINVAL
    assert b"too large" in lib.su2_last_error()
    assert lib.su2_lattice_init(C.byref(l), 1126, 1126, 1126, 1.0, 0.1, 1) == 0
    assert lib.su2_lattice_init(C.byref(l), 1128, 1128, 1128, 1.0, 0.1, 1) == L.SU2_EINVAL
    big = q.Job(1128, 1128, 1128, 3.0)
    assert lib.su2_job_bytes(C.byref(big.c()), 0) == 0


from hypothesis import given, settings, strategies as st  # noqa: E402


@settings(max_examples=300, deadline=None)
@given(st.integers(-3, 1200), st.integers(-3, 1200), st.integers(-3, 1200))
def test_lattice_init_accepts_exactly_the_valid_shapes(Lt, Lx, Ly):
    lib = q.load_library()
    l = L.Su2Lattice()
    rc = lib.su2_lattice_init(C.byref(l), Lt, Lx, Ly, 2.0, 0.1, 1)
    even = all(n >= 2 and n % 2 == 0 for n in (Lt, Lx, Ly))
    if not even:
        assert rc == L.SU2_EODD_EXTENT
    elif Lt * Lx * Ly > 1431633920:
        assert rc == L.SU2_EINVAL
    else:
        assert rc == 0
        assert (l.t, l.x, l.y, l.V, l.d, l.dg, l.dt) == (Lt, Lx, Ly, Lt * Lx * Ly, 3, 4, 12)
        assert lib.su2_field_bytes(C.byref(l), 0) == 96 * l.V
        assert lib.su2_field_bytes(C.byref(l), 1) == 48 * l.V
        assert lib.su2_rng_bytes(C.byref(l)) == 12 * l.V
        assert lib.su2_accept_bytes(C.byref(l)) == 4 * l.V


@settings(max_examples=300, deadline=None)
@given(st.integers(0, 70), st.integers(0, 70), st.integers(0, 70), st.integers(-1, 3), st.integers(-1, 2),
       st.integers(-1, 3), st.integers(-1, 5), st.integers(-1, 5), st.integers(-1, 3), st.integers(-1, 3))
def test_job_bytes_rejects_invalid_jobs(Lt, Lx, Ly, bc, algo, start, n_therm, n_meas, n_cor, hits):
    lib = q.load_library()
    j = q.Job(Lt, Lx, Ly, 3.0, bc=bc, algo=algo, start=start, n_therm=n_therm, n_meas=n_meas, n_cor=n_cor,
              hits=hits)
    valid = (all(n >= 2 and n % 2 == 0 for n in (Lt, Lx, Ly)) and 0 <= bc <= 2 and algo in (0, 1)
             and 0 <= start <= 2 and n_therm >= 1 and n_meas >= 1 and n_cor >= 1
             and (algo != 0 or hits >= 1))
    b64 = lib.su2_job_bytes(C.byref(j.c()), 0)
    b32 = lib.su2_job_bytes(C.byref(j.c()), 1)
    if not valid:
        assert b64 == 0 and b32 == 0
    else:
        V = Lt * Lx * Ly
        # field + RNG + acceptance + partials + two series, each 256-byte aligned
        assert b64 >= 96 * V + 12 * V + 4 * V + 16 * n_meas and b64 % 256 == 0
        assert b32 >= 48 * V + 12 * V + 4 * V and b32 < b64


def test_server_creation_error_comes_back_from_its_thread():
    """The server's thread creates its engine; a configuration the engine
    refuses (before any device call) fails su2_server_create with the
    engine's status and message on the caller's thread, and leaves no
    server behind."""
    from qcd_2d_gpu_amd import engine as E

    lib = q.load_library()
    for cfg, word in ((E.Su2EngineConfig(5, 0, 4, 1, 1 << 20, 0, 0, 0), b"precision"),
                      (E.Su2EngineConfig(0, 0, 4, 1, 1 << 20, 0, 100, 0), b"block")):
        h = C.c_void_p(0x1234)
        assert lib.su2_server_create(C.byref(cfg), C.byref(h)) == L.SU2_EINVAL
        assert h.value is None
        assert word in lib.su2_last_error(), lib.su2_last_error()
