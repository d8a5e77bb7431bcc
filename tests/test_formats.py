"""Drop-in file formats (CPU): the reference's canonical parameter, workflow,
source and receiver files (tests/golden, copied from the reference's par/
directory) and round trips of the model and SU formats."""
import ctypes as C
import os

import numpy as np
import pytest

from common import ROOT, product

G = os.path.join(ROOT, "tests", "golden")


def _lib():
    dn = product()
    L = dn.lib()
    L.dn_par_read.restype = C.c_void_p
    L.dn_par_read.argtypes = [C.c_char_p]
    L.dn_par_free.argtypes = [C.c_void_p]
    for f in ("dn_par_get_int", "dn_par_get_float"):
        getattr(L, f).argtypes = [C.c_void_p, C.c_char_p, C.c_void_p]
    L.dn_par_get_floats.argtypes = [C.c_void_p, C.c_char_p, C.POINTER(C.c_float), C.c_int]
    L.dn_par_get_string.argtypes = [C.c_void_p, C.c_char_p, C.c_char_p, C.c_int]
    L.dn_workflow_read.restype = C.c_void_p
    L.dn_workflow_read.argtypes = [C.c_char_p]
    L.dn_workflow_free.argtypes = [C.c_void_p]
    L.dn_workflow_nstages.argtypes = [C.c_void_p]
    L.dn_workflow_get.argtypes = [C.c_void_p, C.c_int, C.c_char_p, C.POINTER(C.c_float)]
    return L


class Par:
    def __init__(self, path):
        self.L = _lib()
        self.h = self.L.dn_par_read(path.encode())
        assert self.h, self.L.dn_last_error(None)

    def i(self, k):
        v = C.c_int()
        assert self.L.dn_par_get_int(self.h, k.encode(), C.byref(v)) == 0, k
        return v.value

    def f(self, k):
        v = C.c_float()
        assert self.L.dn_par_get_float(self.h, k.encode(), C.byref(v)) == 0, k
        return v.value

    def s(self, k):
        b = C.create_string_buffer(256)
        assert self.L.dn_par_get_string(self.h, k.encode(), b, 256) == 0, k
        return b.value.decode()

    def fl(self, k, n=8):
        a = (C.c_float * n)()
        m = self.L.dn_par_get_floats(self.h, k.encode(), a, n)
        return list(a)[:m]


def test_canonical_inp_positional():
    p = Par(os.path.join(G, "DENISE_SH_canon.inp"))
    assert (p.i("NPROCX"), p.i("NPROCY"), p.i("FDORDER"), p.i("MAXRELERROR")) == (4, 4, 12, 1)
    assert (p.i("NX"), p.i("NY"), p.f("DH"), p.f("TIME"), p.f("DT")) == (64, 64, 25.0, 3.0, np.float32(1e-2))
    assert (p.i("QUELLART"), p.i("QUELLTYP")) == (6, 3)
    assert p.s("SOURCE_FILE") == "./source/source_canon_SH.dat"
    assert p.s("MFILE") == "start/canon_SH_true"
    assert p.i("L") == 4
    np.testing.assert_allclose(p.fl("FL"), [1.0, 11.4657, 2.5295, 28.2239], rtol=1e-6)
    assert (p.i("FREE_SURF"), p.i("FW"), p.f("DAMPING"), p.f("FPML")) == (0, 10, 1000.0, 6.5)
    assert (p.f("REFRECX"), p.f("REFRECY")) == (0.0, 0.0)
    assert (p.f("XREC1"), p.f("YREC1"), p.f("XREC2"), p.f("YREC2")) == (6.0, np.float32(0.2), 93.0, np.float32(0.2))
    assert (p.i("GRADT1"), p.i("GRADT4")) == (1, 5000)
    assert (p.i("INVMAT1"), p.i("INVMAT"), p.i("QUELLTYPB")) == (1, 10, 2)
    assert (p.i("TESTSHOT_START"), p.i("TESTSHOT_END"), p.i("TESTSHOT_INCR")) == (4, 84, 10)
    assert p.s("SEIS_FILE_VY") == "su/DENISE_canon_y.su"
    assert p.i("ITERMAX") == 600
    assert p.s("DATA_DIR_T0") != ""


def test_json_parameters(tmp_path):
    js = tmp_path / "p.json"
    js.write_text('{\n"Comment" : "a comment line, ignored",\n"NX,NY" : "120,40",\n"DH" : "0.25",\n'
                  '"L" : "2",\n"FL1" : "6.25",\n"FL2" : "50.0",\n"REFRECX, REFRECY" : "0.5, 1.5",\n'
                  '"FW" : "-3",\n"MFILE" : "model/start",\n}\n')
    p = Par(str(js))
    assert (p.i("NX"), p.i("NY"), p.f("DH"), p.i("L")) == (120, 40, 0.25, 2)
    assert p.fl("FL") == [6.25, 50.0]
    assert (p.f("REFRECX"), p.f("REFRECY")) == (0.5, 1.5)
    assert p.i("FW") == 0  # read_par.c:266: FW < 0 -> 0
    assert p.s("MFILE") == "model/start"


def test_canonical_workflow():
    L = _lib()
    w = L.dn_workflow_read(os.path.join(G, "FWI_workflow_SH_canon.inp").encode())
    assert w
    assert L.dn_workflow_nstages(w) == 3
    v = C.c_float()
    got = []
    for st in (1, 2, 3):
        assert L.dn_workflow_get(w, st, b"FC_END", C.byref(v)) == 0
        got.append(v.value)
    assert got == [4.0, 8.0, 13.0]
    assert L.dn_workflow_get(w, 2, b"SCALERHO", C.byref(v)) == 0 and v.value == 0.5
    assert L.dn_workflow_get(w, 1, b"SPATFILTER", C.byref(v)) == 0 and v.value == 4.0
    L.dn_workflow_free(w)


def test_canonical_sources_and_receivers():
    L = _lib()
    src = np.zeros((200, 8), np.float32)
    n = C.c_int()
    L.dn_read_sources.argtypes = [C.c_char_p, C.c_int, C.c_void_p, C.POINTER(C.c_int)]
    assert L.dn_read_sources(os.path.join(G, "source_canon_SH.dat").encode(), 200, src.ctypes.data,
                             C.byref(n)) == 0
    assert n.value == 84
    # columns x z y tshift fc amp azimuth type -> srcpos x, y, 0, tshift, fc, amp, azimuth, type
    np.testing.assert_array_equal(src[0], [250.0, 250.0, 0.0, 0.0, 10.0, 1.0, 0.0, 3.0])
    ri = np.zeros(200, np.int32)
    rj = np.zeros(200, np.int32)
    L.dn_read_receivers.argtypes = [C.c_char_p, C.c_float, C.c_float, C.c_float, C.c_int, C.c_void_p, C.c_void_p,
                                    C.POINTER(C.c_int)]
    assert L.dn_read_receivers(os.path.join(G, "receiver_canon_SH.dat").encode(), 25.0, 0.0, 0.0, 200,
                               ri.ctypes.data, rj.ctypes.data, C.byref(n)) == 0
    xy = np.loadtxt(os.path.join(G, "receiver_canon_SH.dat"))
    exp = np.floor(xy / np.float32(25.0) + 0.5).astype(int)
    # first occurrences of each grid point, x index > 0
    seen, keep = set(), []
    for a, b in exp:
        if (a, b) not in seen and a > 0:
            keep.append((a, b))
        seen.add((a, b))
    assert n.value == len(keep)
    assert list(zip(ri[:n.value], rj[:n.value])) == keep


def test_model_file_round_trip_is_x_major(tmp_path):
    L = _lib()
    nx, ny = 7, 5
    a = np.arange(nx * ny, dtype=np.float32).reshape(ny, nx)
    path = str(tmp_path / "m.vs").encode()
    L.dn_write_model_file.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_void_p]
    L.dn_read_model_file.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_void_p]
    assert L.dn_write_model_file(path, nx, ny, a.ctypes.data) == 0
    raw = np.fromfile(path.decode(), np.float32)
    np.testing.assert_array_equal(raw, a.T.ravel())  # all NY values of column 1 first (readmod.c:82-85)
    b = np.zeros_like(a)
    assert L.dn_read_model_file(path, nx, ny, b.ctypes.data) == 0
    np.testing.assert_array_equal(a, b)


def test_su_round_trip_and_header(tmp_path):
    L = _lib()
    ntr, ns = 4, 30
    data = np.random.default_rng(0).standard_normal((ntr, ns)).astype(np.float32)
    ri = np.array([10, 20, 30, 40], np.int32)
    rj = np.array([2, 2, 2, 2], np.int32)
    path = str(tmp_path / "d.su").encode()
    f = C.c_float
    L.dn_write_su.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, f, f, f, f, f, f,
                              C.c_int]
    assert L.dn_write_su(path, ntr, ns, data.ctypes.data, ri.ctypes.data, rj.ctypes.data, 0.25, 0.0, 0.0, 5.0, 0.5,
                         1e-4, 1) == 0
    raw = open(path.decode(), "rb").read()
    assert len(raw) == ntr * (240 + 4 * ns)
    hdr = np.frombuffer(raw[:240], np.uint8)
    i32 = lambda off: int(np.frombuffer(hdr[off:off + 4].tobytes(), np.int32)[0])
    u16 = lambda off: int(np.frombuffer(hdr[off:off + 2].tobytes(), np.uint16)[0])
    i16 = lambda off: int(np.frombuffer(hdr[off:off + 2].tobytes(), np.int16)[0])
    assert u16(114) == ns and u16(116) == 100          # ns, dt in microseconds
    assert i16(28) == 1 and i16(68) == -1000 and i16(70) == -1000
    assert i32(72) == 5000 and i32(80) == 2500          # sx, gx in mm
    assert i32(36) == int(np.floor(np.sqrt((5.0 - 2.5) ** 2 + (0.5 - 0.5) ** 2) * 1000 + 0.5))
    tr0 = np.frombuffer(raw[240:240 + 4 * ns], np.float32)
    np.testing.assert_array_equal(tr0, data[0])
    back = np.zeros_like(data)
    L.dn_read_su.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_void_p]
    assert L.dn_read_su(path, ntr, ns, back.ctypes.data) == 0
    exp = data.copy()
    exp[:, 0] = 0.0   # inseis.c:76
    np.testing.assert_array_equal(back, exp)
