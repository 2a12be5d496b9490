"""tb8gen_amd -- MI355X king-slice retrograde engine (Python side of the C ABI).

The engine is libtbg.so (HIP kernels + the C ABI declared in include/tbg.h);
this module binds it with ctypes and mirrors the reference generator's
driver (syzygy1/tb8gen src/tb8gen.c:1245-1453): a material is generated
after every smaller material its captures reach, and the per-ply count
lines come out in the reference's format.  There is no CPU fallback: if the
library or a HIP device is missing, every call raises.
"""
import ctypes
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
LIB_PATH = os.path.join(PKG, "libtbg.so")
DRIVER_PATH = os.path.join(PKG, "tb8gen")

FORM_PLAIN, FORM_BATCHED, FORM_GRAPH = 0, 1, 2
PLANE_WINS, PLANE_W, PLANE_L = 0, 1, 2
SUB_LOSS, SUB_BLOSS, SUB_CWIN, SUB_WIN = 0, 1, 2, 3
NSLICES = 462
DRAW_RULE = 100

# the exported entry points of include/tbg.h
ABI_SYMBOLS = [
    "tbg_init", "tbg_strerror", "tbg_create", "tbg_create_synth", "tbg_create_synth_ex",
    "tbg_destroy",
    "tbg_get_info", "tbg_resolve_captures", "tbg_init_illegal_mate", "tbg_step_L",
    "tbg_step_W", "tbg_step_L_plain", "tbg_step_W_plain", "tbg_end_ply", "tbg_generate",
    "tbg_get_codes", "tbg_get_bitmap", "tbg_get_sub_bitmap", "tbg_get_timing",
    "tbg_set_profiling", "tbg_host_selftest", "tbg_set_scatter", "tbg_generate_batch",
    "tbg_emit_lines", "tbg_generate_batch_noalloc", "tbg_server_create", "tbg_server_destroy", "tbg_server_submit",
    "tbg_server_run", "tbg_server_result", "tbg_server_codes", "tbg_server_lines",
    "tbg_server_serve", "tbg_server_close", "tbg_job_poll", "tbg_server_codes_part",
    "tbg_server_lines_part", "tbg_king_slice", "tbg_slice_kings",
    "tbg_set_geometry", "tbg_get_geometry", "tbg_autotune", "tbg_file_write", "tbg_file_read",
    "tbg_checkpoint", "tbg_restore", "tbg_continue", "tbg_export_plies", "tbg_rtbw_name",
    "tbg_write_rtbw",
    "tbg_registry_create", "tbg_registry_get", "tbg_registry_count", "tbg_registry_entry",
    "tbg_registry_destroy",
]
SCATTER_GLOBAL, SCATTER_BLOCK, SCATTER_COMBINE = 0, 1, 2
SYNTH_CODES = 1
NOALLOC_MAX_BATCH = 256


class TbgError(RuntimeError):
    pass


class Info(ctypes.Structure):
    _fields_ = [("npieces", ctypes.c_int), ("pt", ctypes.c_uint8 * 8), ("numsets", ctypes.c_int),
                ("mult", ctypes.c_int * 6), ("slice_size", ctypes.c_uint64),
                ("slice_words", ctypes.c_uint64), ("capt_size", ctypes.c_uint64 * 6),
                ("sub_words", ctypes.c_uint64 * 2), ("sub_offset", ctypes.c_uint64 * 6),
                ("device_bytes", ctypes.c_uint64), ("synth_slice", ctypes.c_int)]


class Counts(ctypes.Structure):
    _fields_ = [("broken", ctypes.c_uint64 * 2), ("mates", ctypes.c_uint64 * 2),
                ("nplies", ctypes.c_int)]


class Timing(ctypes.Structure):
    _fields_ = [("ms_total", ctypes.c_double), ("ms_pred", ctypes.c_double),
                ("ms_verify", ctypes.c_double), ("ms_sweep", ctypes.c_double),
                ("ms_uncap", ctypes.c_double), ("ms_init", ctypes.c_double),
                ("ms_resolve", ctypes.c_double), ("launches", ctypes.c_uint64),
                ("half_steps", ctypes.c_uint64), ("n_pred", ctypes.c_uint64),
                ("n_verify", ctypes.c_uint64), ("n_sweep", ctypes.c_uint64),
                ("n_uncap", ctypes.c_uint64), ("n_init", ctypes.c_uint64),
                ("n_resolve", ctypes.c_uint64), ("ms_xr", ctypes.c_double),
                ("n_xr", ctypes.c_uint64)]


class Geometry(ctypes.Structure):
    _fields_ = [("verify_unit_words", ctypes.c_int), ("sweep_unit_words", ctypes.c_int),
                ("pred_wg_per_cu", ctypes.c_int), ("uncap_wg_per_cu", ctypes.c_int),
                ("verify_wg_per_cu", ctypes.c_int), ("sweep_wg_per_cu", ctypes.c_int),
                ("scatter", ctypes.c_int)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class JobSpec(ctypes.Structure):
    _fields_ = [("material", ctypes.c_char_p), ("slice", ctypes.c_int), ("seed", ctypes.c_uint64),
                ("arrival_ms", ctypes.c_double), ("nslices", ctypes.c_int)]


class JobResult(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int), ("nplies", ctypes.c_int), ("arrival_ms", ctypes.c_double),
                ("admit_ms", ctypes.c_double), ("done_ms", ctypes.c_double),
                ("positions", ctypes.c_uint64), ("device_bytes", ctypes.c_uint64),
                ("half_steps", ctypes.c_uint64), ("nslices", ctypes.c_int),
                ("spills", ctypes.c_int)]


class ServerStats(ctypes.Structure):
    _fields_ = [("wall_ms", ctypes.c_double), ("jobs_done", ctypes.c_uint64),
                ("jobs_rejected", ctypes.c_uint64), ("mean_resident_bytes", ctypes.c_double),
                ("peak_resident_bytes", ctypes.c_uint64), ("peak_active", ctypes.c_int),
                ("plies", ctypes.c_uint64), ("position_side_plies", ctypes.c_double),
                ("spills", ctypes.c_uint64), ("restores", ctypes.c_uint64),
                ("spill_ms", ctypes.c_double)]


LINE_FN = ctypes.CFUNCTYPE(None, ctypes.c_char_p, ctypes.c_void_p)

_lib = None
_inited = False


def lib():
    """Load libtbg.so (built in-tree by tb8gen_amd.build); raises if absent."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise TbgError(f"{LIB_PATH} is not built (run python -m tb8gen_amd.build)")
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        L.tbg_strerror.restype = ctypes.c_char_p
        L.tbg_create.argtypes = [ctypes.POINTER(vp), ctypes.c_char_p]
        L.tbg_create_synth.argtypes = [ctypes.POINTER(vp), ctypes.c_char_p, ctypes.c_int,
                                       ctypes.c_uint64]
        L.tbg_create_synth_ex.argtypes = [ctypes.POINTER(vp), ctypes.c_char_p, ctypes.c_int,
                                          ctypes.c_uint64, ctypes.c_uint]
        L.tbg_destroy.argtypes = [vp]
        L.tbg_destroy.restype = None
        L.tbg_get_info.argtypes = [vp, ctypes.POINTER(Info)]
        L.tbg_resolve_captures.argtypes = [vp, ctypes.POINTER(vp), ctypes.c_int]
        L.tbg_init_illegal_mate.argtypes = [vp, ctypes.POINTER(Counts)]
        for f in ("tbg_step_L", "tbg_step_W", "tbg_step_L_plain", "tbg_step_W_plain"):
            getattr(L, f).argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.POINTER(ctypes.c_uint64)]
        L.tbg_end_ply.argtypes = [vp]
        L.tbg_generate.argtypes = [vp, ctypes.c_int, LINE_FN, vp, ctypes.POINTER(Counts)]
        L.tbg_get_codes.argtypes = [vp, ctypes.c_int, ctypes.c_int,
                                    ctypes.POINTER(ctypes.c_uint16), ctypes.c_uint64]
        L.tbg_get_bitmap.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     ctypes.POINTER(ctypes.c_uint64), ctypes.c_uint64]
        L.tbg_get_sub_bitmap.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         ctypes.POINTER(ctypes.c_uint64), ctypes.c_uint64]
        L.tbg_get_timing.argtypes = [vp, ctypes.POINTER(Timing)]
        L.tbg_set_profiling.argtypes = [vp, ctypes.c_int]
        L.tbg_set_scatter.argtypes = [vp, ctypes.c_int]
        L.tbg_generate_batch.argtypes = [ctypes.POINTER(vp), ctypes.c_int, ctypes.c_int,
                                         ctypes.POINTER(Counts)]
        L.tbg_emit_lines.argtypes = [vp, LINE_FN, vp]
        L.tbg_generate_batch_noalloc.argtypes = [ctypes.POINTER(vp), ctypes.c_int,
                                                 ctypes.POINTER(Counts)]
        L.tbg_server_create.argtypes = [ctypes.POINTER(vp), ctypes.c_uint64, ctypes.c_int]
        L.tbg_server_destroy.argtypes = [vp]
        L.tbg_server_destroy.restype = None
        L.tbg_server_submit.argtypes = [vp, ctypes.POINTER(JobSpec), ctypes.POINTER(ctypes.c_uint64)]
        L.tbg_server_run.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ServerStats)]
        L.tbg_server_serve.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ServerStats)]
        L.tbg_server_close.argtypes = [vp]
        L.tbg_job_poll.argtypes = [vp, ctypes.c_uint64, ctypes.POINTER(JobResult)]
        L.tbg_server_codes_part.argtypes = [vp, ctypes.c_uint64, ctypes.c_int, ctypes.c_int,
                                            ctypes.POINTER(ctypes.c_uint16), ctypes.c_uint64]
        L.tbg_server_lines_part.argtypes = [vp, ctypes.c_uint64, ctypes.c_int, LINE_FN, vp]
        L.tbg_server_result.argtypes = [vp, ctypes.c_uint64, ctypes.POINTER(JobResult)]
        L.tbg_server_codes.argtypes = [vp, ctypes.c_uint64, ctypes.c_int,
                                       ctypes.POINTER(ctypes.c_uint16), ctypes.c_uint64]
        L.tbg_server_lines.argtypes = [vp, ctypes.c_uint64, LINE_FN, vp]
        L.tbg_host_selftest.argtypes = [ctypes.c_char_p, ctypes.c_int]
        L.tbg_king_slice.argtypes = [ctypes.c_int, ctypes.c_int]
        L.tbg_slice_kings.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                      ctypes.POINTER(ctypes.c_int)]
        L.tbg_set_geometry.argtypes = [vp, ctypes.POINTER(Geometry)]
        L.tbg_get_geometry.argtypes = [vp, ctypes.POINTER(Geometry)]
        L.tbg_autotune.argtypes = [vp, ctypes.c_int, ctypes.c_char_p, ctypes.POINTER(Geometry),
                                   ctypes.POINTER(ctypes.c_double)]
        L.tbg_file_write.argtypes = [ctypes.c_char_p, vp, ctypes.c_uint64]
        L.tbg_file_read.argtypes = [ctypes.c_char_p, vp, ctypes.c_uint64]
        L.tbg_checkpoint.argtypes = [vp, ctypes.c_char_p]
        L.tbg_restore.argtypes = [vp, ctypes.c_char_p]
        L.tbg_continue.argtypes = [vp, ctypes.c_int, ctypes.c_int, LINE_FN, vp,
                                   ctypes.POINTER(Counts)]
        L.tbg_export_plies.argtypes = [vp, ctypes.c_char_p]
        L.tbg_rtbw_name.argtypes = [vp, ctypes.c_char_p, ctypes.c_size_t]
        L.tbg_write_rtbw.argtypes = [vp, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_size_t]
        L.tbg_registry_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int]
        L.tbg_registry_get.argtypes = [vp, ctypes.c_char_p, ctypes.POINTER(vp)]
        L.tbg_registry_count.argtypes = [vp]
        L.tbg_registry_entry.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_char_p),
                                         ctypes.POINTER(vp)]
        L.tbg_registry_destroy.argtypes = [vp]
        L.tbg_registry_destroy.restype = None
        _lib = L
    return _lib


def check(r, what=""):
    if r != 0:
        msg = lib().tbg_strerror(r).decode()
        raise TbgError(f"{what}: {msg} ({r})")


def init(device=0):
    global _inited
    if not _inited:
        check(lib().tbg_init(device), "tbg_init")
        _inited = True


def host_selftest():
    buf = ctypes.create_string_buffer(256)
    r = lib().tbg_host_selftest(buf, 256)
    return r, buf.value.decode()


def king_slice(wk, bk):
    """Slice number of white king wk and black king bk (-1: no slice)."""
    return lib().tbg_king_slice(int(wk), int(bk))


def slice_kings(s):
    """(white king, black king) squares numbering slice s."""
    wk, bk = ctypes.c_int(), ctypes.c_int()
    check(lib().tbg_slice_kings(int(s), ctypes.byref(wk), ctypes.byref(bk)), f"slice {s}")
    return wk.value, bk.value


class Table:
    """Device-resident state of one material (a tbg_ctx)."""

    def __init__(self, material, synth_slice=None, seed=0, ptr=None, keep_codes=False):
        """ptr: wrap a context owned elsewhere (a NativeRegistry's); close()
        then leaves it alone.  keep_codes: a synthetic slice also keeps its
        per-position codes (whole tables always do)."""
        init()
        self.material = material
        self.owned = ptr is None
        self.ptr = ctypes.c_void_p(ptr) if ptr is not None else ctypes.c_void_p()
        if ptr is not None:
            pass
        elif synth_slice is None:
            check(lib().tbg_create(ctypes.byref(self.ptr), material.encode()), material)
        else:
            check(lib().tbg_create_synth_ex(ctypes.byref(self.ptr), material.encode(),
                                            int(synth_slice), int(seed),
                                            SYNTH_CODES if keep_codes else 0), material)
        self.synth_slice = synth_slice
        self.lines = []
        self.counts = Counts()
        info = Info()
        check(lib().tbg_get_info(self.ptr, ctypes.byref(info)), "info")
        self.info = info
        self.size = info.slice_size
        self.words = info.slice_words

    def close(self):
        if self.ptr and self.owned:
            lib().tbg_destroy(self.ptr)
        self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def resolve(self, subs=()):
        arr = (ctypes.c_void_p * max(1, len(subs)))(*[s.ptr.value for s in subs])
        check(lib().tbg_resolve_captures(self.ptr, arr, len(subs)), f"resolve {self.material}")

    def generate(self, form=FORM_BATCHED):
        self.lines = []

        def cb(line, _user):
            self.lines.append(line.decode())

        fn = LINE_FN(cb)
        check(lib().tbg_generate(self.ptr, form, fn, None, ctypes.byref(self.counts)),
              f"generate {self.material}")
        return self.lines

    def continue_(self, form=FORM_BATCHED, stop_at=0):
        """Run the ply loop on from the current state (after init_illegal_mate,
        restore or an earlier stop), to the fixpoint or to ply stop_at; returns
        the count lines of the plies run so far."""
        self.lines = []

        def cb(line, _user):
            self.lines.append(line.decode())

        fn = LINE_FN(cb)
        check(lib().tbg_continue(self.ptr, form, int(stop_at), fn, None,
                                 ctypes.byref(self.counts)), f"continue {self.material}")
        return self.lines

    def checkpoint(self, root):
        check(lib().tbg_checkpoint(self.ptr, str(root).encode()), f"checkpoint {self.material}")

    def restore(self, root):
        check(lib().tbg_restore(self.ptr, str(root).encode()), f"restore {self.material}")

    def export_plies(self, root):
        check(lib().tbg_export_plies(self.ptr, str(root).encode()), f"export {self.material}")

    def rtbw_name(self):
        """Name of the table's WDL file in the probing code's colour convention."""
        buf = ctypes.create_string_buffer(32)
        check(lib().tbg_rtbw_name(self.ptr, buf, 32), f"rtbw name {self.material}")
        return buf.value.decode()

    def write_rtbw(self, directory):
        """Writes the finished table as <directory>/<name>.rtbw (Syzygy WDL
        format); returns the path."""
        buf = ctypes.create_string_buffer(4096)
        check(lib().tbg_write_rtbw(self.ptr, str(directory).encode(), buf, 4096),
              f"write rtbw {self.material}")
        return buf.value.decode()

    def emit_lines(self):
        self.lines = []

        def cb(line, _user):
            self.lines.append(line.decode())

        fn = LINE_FN(cb)
        check(lib().tbg_emit_lines(self.ptr, fn, None), f"lines {self.material}")
        return self.lines

    def step(self, kind, stm, n, more, plain=False):
        """One host-driven half-step (kind 'L' or 'W'); returns the count."""
        cnt = ctypes.c_uint64()
        f = {("L", False): "tbg_step_L", ("W", False): "tbg_step_W",
             ("L", True): "tbg_step_L_plain", ("W", True): "tbg_step_W_plain"}[(kind, plain)]
        check(getattr(lib(), f)(self.ptr, stm, n, 1 if more else 0, ctypes.byref(cnt)), f)
        return cnt.value

    def init_illegal_mate(self):
        c = Counts()
        check(lib().tbg_init_illegal_mate(self.ptr, ctypes.byref(c)), "init")
        return c

    def end_ply(self):
        check(lib().tbg_end_ply(self.ptr), "end_ply")

    def codes(self, stm, slice_=None):
        """uint16 codes [462, size] (or one slice)."""
        slices = [slice_] if slice_ is not None else (
            [self.synth_slice] if self.synth_slice is not None else range(NSLICES))
        out = np.zeros((len(slices), self.size), dtype=np.uint16)
        for i, s in enumerate(slices):
            check(lib().tbg_get_codes(self.ptr, stm, s,
                                      out[i].ctypes.data_as(ctypes.POINTER(ctypes.c_uint16)),
                                      self.size), "codes")
        return out

    def bitmap(self, plane, stm, s):
        out = np.zeros(self.words, dtype=np.uint64)
        check(lib().tbg_get_bitmap(self.ptr, plane, stm, s,
                                   out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                   self.words), "bitmap")
        return out

    def sub_bitmap(self, kind, colour, s):
        w = self.info.sub_words[colour]
        out = np.zeros(w, dtype=np.uint64)
        check(lib().tbg_get_sub_bitmap(self.ptr, kind, colour, s,
                                       out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), w),
              "sub bitmap")
        return out

    def set_scatter(self, mode):
        check(lib().tbg_set_scatter(self.ptr, int(mode)), "scatter")

    def set_profiling(self, on):
        check(lib().tbg_set_profiling(self.ptr, 1 if on else 0), "profiling")

    def geometry(self):
        g = Geometry()
        check(lib().tbg_get_geometry(self.ptr, ctypes.byref(g)), "geometry")
        return g

    def set_geometry(self, g):
        check(lib().tbg_set_geometry(self.ptr, ctypes.byref(g)), "geometry")

    def autotune(self, max_plies=0, cache_path=None):
        """Pick the launch geometry by measurement (captures resolved first);
        returns (Geometry, ms of the chosen trial, -1 on a cache hit).  The
        table must be generated again afterwards."""
        g = Geometry()
        ms = ctypes.c_double()
        check(lib().tbg_autotune(self.ptr, int(max_plies),
                                 cache_path.encode() if cache_path else None,
                                 ctypes.byref(g), ctypes.byref(ms)), "autotune")
        return g, ms.value

    def timing(self):
        t = Timing()
        check(lib().tbg_get_timing(self.ptr, ctypes.byref(t)), "timing")
        return t


def file_write(path, data):
    """Write bytes (a numpy array or bytes) as a slice file in the reference's
    zstd frame format."""
    import numpy as np
    a = np.ascontiguousarray(np.frombuffer(bytes(data), dtype=np.uint8)
                             if isinstance(data, (bytes, bytearray)) else data)
    check(lib().tbg_file_write(str(path).encode(), a.ctypes.data_as(ctypes.c_void_p), a.nbytes),
          f"write {path}")


def file_read(path, nbytes):
    """Read a slice file of nbytes decompressed bytes; returns a uint8 array."""
    import numpy as np
    out = np.zeros(nbytes, dtype=np.uint8)
    check(lib().tbg_file_read(str(path).encode(), out.ctypes.data_as(ctypes.c_void_p), nbytes),
          f"read {path}")
    return out


def default_geometry():
    g = Geometry()
    check(lib().tbg_get_geometry(None, ctypes.byref(g)), "geometry")
    return g


def set_default_geometry(g):
    """Geometry of contexts and servers created from now on."""
    check(lib().tbg_set_geometry(None, ctypes.byref(g)), "geometry")


def generate_batch(tables, form=FORM_GRAPH):
    """Generate several independent materials together (one launch per phase
    covers all of them); returns each table's count lines."""
    arr = (ctypes.c_void_p * len(tables))(*[t.ptr.value for t in tables])
    counts = (Counts * len(tables))()
    check(lib().tbg_generate_batch(arr, len(tables), form, counts), "generate_batch")
    for t, c in zip(tables, counts):
        t.counts = c
        t.emit_lines()
    return [t.lines for t in tables]


def generate_batch_noalloc(tables):
    """generate_batch in the batched form without any device allocation
    (tbg_generate_batch_noalloc; at most NOALLOC_MAX_BATCH tables)."""
    arr = (ctypes.c_void_p * len(tables))(*[t.ptr.value for t in tables])
    counts = (Counts * len(tables))()
    check(lib().tbg_generate_batch_noalloc(arr, len(tables), counts), "generate_batch_noalloc")
    for t, c in zip(tables, counts):
        t.counts = c
        t.emit_lines()
    return [t.lines for t in tables]


class Server:
    """Continuous-batching server of synthetic king-slice jobs (libtbg).
    Jobs may be submitted before run()/serve() or, with serve() running on
    another thread, while it runs; poll() tells whether one is done."""

    def __init__(self, budget_bytes, max_active=64):
        init()
        self.ptr = ctypes.c_void_p()
        check(lib().tbg_server_create(ctypes.byref(self.ptr), int(budget_bytes), int(max_active)),
              "server")
        self.ids = []

    def submit(self, material, slice_, seed, arrival_ms=0.0, nslices=1):
        """arrival_ms < 0: the moment of submission (live serving)."""
        spec = JobSpec(material.encode(), int(slice_), int(seed), float(arrival_ms), int(nslices))
        jid = ctypes.c_uint64()
        check(lib().tbg_server_submit(self.ptr, ctypes.byref(spec), ctypes.byref(jid)), "submit")
        self.ids.append(jid.value)
        return jid.value

    def run(self, form=FORM_BATCHED):
        """Serve every job submitted so far (closes the server to new jobs)."""
        st = ServerStats()
        check(lib().tbg_server_run(self.ptr, form, ctypes.byref(st)), "server run")
        return st

    def serve(self, form=FORM_BATCHED):
        """The serving loop; returns once close() was called and every job
        is done.  ctypes releases the interpreter lock for its duration, so
        other threads submit and poll meanwhile."""
        st = ServerStats()
        check(lib().tbg_server_serve(self.ptr, form, ctypes.byref(st)), "server serve")
        return st

    def close_submissions(self):
        check(lib().tbg_server_close(self.ptr), "server close")

    def poll(self, jid):
        """The job's result once it is done, else None."""
        r = JobResult()
        rc = lib().tbg_job_poll(self.ptr, jid, ctypes.byref(r))
        if rc == 1:  # TBG_PENDING
            return None
        check(rc, "poll")
        return r

    def result(self, jid):
        r = JobResult()
        check(lib().tbg_server_result(self.ptr, jid, ctypes.byref(r)), "result")
        return r

    def codes(self, jid, stm, part=0):
        r = self.result(jid)
        out = np.zeros(r.positions, dtype=np.uint16)
        check(lib().tbg_server_codes_part(self.ptr, jid, int(part), stm,
                                          out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16)),
                                          r.positions), "codes")
        return out

    def lines(self, jid, part=0):
        out = []

        def cb(line, _u):
            out.append(line.decode())

        fn = LINE_FN(cb)
        check(lib().tbg_server_lines_part(self.ptr, jid, int(part), fn, None), "lines")
        return out

    def close(self):
        if self.ptr:
            lib().tbg_server_destroy(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_ORDER = "KQRBN"


def _side_key(side):
    return (len(side), [-_ORDER.index(ch) for ch in side])


def canonical(material):
    """Colour orientation used for generated smaller materials: the side with
    more pieces, then the stronger piece string, is white."""
    w, b = material.split("v")
    w = "".join(sorted(w, key=_ORDER.index))
    b = "".join(sorted(b, key=_ORDER.index))
    return f"{w}v{b}" if _side_key(w) >= _side_key(b) else f"{b}v{w}"


def sub_materials(material):
    """Materials one capture away (excluding a bare pair of kings)."""
    w, b = material.split("v")
    out = []
    for side, other, white in ((w, b, True), (b, w, False)):
        for i, ch in enumerate(side):
            if ch == "K":
                continue
            rest = side[:i] + side[i + 1:]
            m = f"{rest}v{other}" if white else f"{other}v{rest}"
            if len(m) - 1 > 2:
                c = canonical(m)
                if c not in out:
                    out.append(c)
    return out


class NativeRegistry:
    """The library's material scheduler (tbg_registry_*): a material and every
    smaller material its captures reach, generated smallest first, one batch
    per piece count, kept resident."""

    def __init__(self, form=FORM_GRAPH):
        init()
        self.ptr = ctypes.c_void_p()
        check(lib().tbg_registry_create(ctypes.byref(self.ptr), form), "registry")

    def get(self, material):
        c = ctypes.c_void_p()
        check(lib().tbg_registry_get(self.ptr, material.encode(), ctypes.byref(c)), material)
        t = Table(material, ptr=c.value)
        t.emit_lines()
        return t

    def count(self):
        return lib().tbg_registry_count(self.ptr)

    def entries(self):
        """(material, Table) of every resident table, smallest first."""
        out = []
        for i in range(self.count()):
            name, c = ctypes.c_char_p(), ctypes.c_void_p()
            check(lib().tbg_registry_entry(self.ptr, i, ctypes.byref(name), ctypes.byref(c)),
                  "registry entry")
            out.append((name.value.decode(), Table(name.value.decode(), ptr=c.value)))
        return out

    @property
    def tables(self):
        """{material: Table} of every resident table."""
        return dict(self.entries())

    def close(self):
        if self.ptr:
            lib().tbg_registry_destroy(self.ptr)
            self.ptr = ctypes.c_void_p()
