"""Python handles on the C++ host layer (include/tti_host.h): the simulator2SetUp and simulator3SetUp drivers and
the snapshot store.

The drivers (simulator2/simulator2SetUp.cu, simulator3/simulator3SetUp.cu) run entirely in C++ on the null stream;
this module only creates them, starts their passes and copies volumes back for inspection.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import TTIError, check, lib


class Simulator2:
    """simulator2SetUp(npoints, npml, niter, xcorr_step, nzones, ranks={1,1,1}, 4, 4, 32, fwd_only, comm)."""

    def __init__(self, npoints: int, npml: int, niter: int, xcorr_step: int, nzones=(1, 1, 1), fwd_only=False,
                 bit_per_float: float = 32.0):
        # 0 < bit_per_float < 32: snapshots stored ZFP-coded at that rate (fixed-rate mode)
        z = (C.c_int * 3)(*nzones)
        h = lib().tti_sim2_create_ex(npoints, npml, niter, xcorr_step, z, int(fwd_only), float(bit_per_float))
        if not h:
            raise TTIError(f"simulator2SetUp: {lib().tti_host_last_error().decode()}")
        self._h = C.c_void_p(h)

    def close(self) -> None:
        if self._h:
            lib().tti_sim2_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_medium(self, vp, eps=0.0, delta=0.0, vs=0.0, theta=0.0, phi=0.0, rho=1.0) -> None:
        """Constant medium in every sub-volume. The time step stays the constructor's (2.70e-4 s, the reference's
        rule for vp = 1500 m/s and 1 m cells): the scheme is stable while the fastest qP speed (vp sqrt(1 + 2 eps)
        for eps >= delta) stays below 1664 m/s."""
        check(lib().tti_sim2_set_medium(self._h, vp, eps, delta, vs, theta, phi, rho), "set_medium")

    def set_free_surface(self, on: bool) -> None:
        check(lib().tti_sim2_set_free_surface(self._h, int(on)), "set_free_surface")

    def set_fused(self, on: bool) -> None:
        """Fused time steps on one zone (default on); off keeps the reference's three-kernel steps."""
        check(lib().tti_sim2_set_fused(self._h, int(on)), "set_fused")

    def set_update_born(self, on: bool) -> None:
        """kernel_base's update_born: the time-update loops also advance the Born fields "field11".."field41" and
        the Born memory variables "Model81_m", "Model91_m" (allocated zero here). Set before the forward pass."""
        check(lib().tti_sim2_set_update_born(self._h, int(on)), "set_update_born")

    @property
    def fused_active(self) -> bool:
        r = lib().tti_sim2_fused_active(self._h)
        if r < 0:
            check(r, "fused_active")
        return r == 1

    def set_snapshot_store(self, scratch: str = ".", max_mem_steps: int = -1) -> None:
        """Snapshot steps beyond max_mem_steps (or the host-RAM budget) go to a scratch file; before forward()."""
        check(lib().tti_sim2_set_snapshot_store(self._h, scratch.encode(), max_mem_steps), "set_snapshot_store")

    def execute(self, has_src_p=True, has_qp=False) -> None:
        check(lib().tti_sim2_execute(self._h, int(has_src_p), int(has_qp)), "execute")

    def forward(self, has_src_p=True) -> None:
        check(lib().tti_sim2_forward(self._h, int(has_src_p)), "forward")

    def adjoint(self) -> None:
        check(lib().tti_sim2_adjoint(self._h), "adjoint")

    @property
    def num_subs(self) -> int:
        return lib().tti_sim2_num_subs(self._h)

    @property
    def dt(self) -> float:
        return lib().tti_sim2_dt(self._h)

    def info(self, name: str, isub: int) -> dict:
        a = (C.c_int * 12)()
        check(lib().tti_sim2_volume_info(self._h, name.encode(), isub, a), "volume_info")
        v = list(a)
        return {"offset": v[0:3], "n": v[3:6], "ntot": v[6:9], "ibeg": v[9:12]}

    def read(self, name: str, isub: int) -> np.ndarray:
        """Whole allocation of a volume, shape (ntot_z, ntot_y, ntot_x)."""
        nt = self.info(name, isub)["ntot"]
        out = np.empty((nt[2], nt[1], nt[0]), dtype=np.float32)
        check(lib().tti_sim2_read(self._h, name.encode(), isub, out.ctypes.data), "read")
        return out

    def write(self, name: str, isub: int, a: np.ndarray) -> None:
        """Whole allocation of a volume from `a` (shape (ntot_z, ntot_y, ntot_x), converted to float32)."""
        nt = self.info(name, isub)["ntot"]
        a = np.ascontiguousarray(a, dtype=np.float32)
        if a.shape != (nt[2], nt[1], nt[0]):
            raise ValueError(f"write {name}: shape {a.shape}, the volume is {(nt[2], nt[1], nt[0])}")
        check(lib().tti_sim2_write(self._h, name.encode(), isub, a.ctypes.data, a.size), "write")

    def interior(self, name: str, isub: int) -> np.ndarray:
        i = self.info(name, isub)
        a = self.read(name, isub)
        (bx, by, bz), (nx, ny, nz) = i["ibeg"], i["n"]
        return a[bz : bz + nz, by : by + ny, bx : bx + nx]

    def assemble(self, name: str) -> np.ndarray:
        """Interior of `name` over the whole grid, assembled from the sub-volumes."""
        parts = [(self.info(name, s), self.interior(name, s)) for s in range(self.num_subs)]
        shape = [max(i["offset"][d] + i["n"][d] for i, _ in parts) for d in range(3)]
        out = np.zeros(shape[::-1], dtype=np.float32)
        for i, a in parts:
            (ox, oy, oz), (nx, ny, nz) = i["offset"], i["n"]
            out[oz : oz + nz, oy : oy + ny, ox : ox + nx] = a
        return out

    def timings(self) -> dict:
        t = (C.c_double * 5)()
        check(lib().tti_sim2_timings(self._h, t), "timings")
        return dict(zip(("forward", "adjoint", "velocity", "stress", "halo"), list(t)))


class Simulator3:
    """simulator3SetUp(npoints, npml, niter, xcorr_step, nzones, ranks={1,1,1}, 4, 4, 32, fwd_only, comm): the
    staggered-grid elastic TTI driver. Volumes: f1..f21, M1..M22, d4..d9, p, g1, g2, snap_f1..snap_f9, zprime."""

    def __init__(self, npoints: int, npml: int, niter: int, xcorr_step: int, nzones=(1, 1, 1), fwd_only=False,
                 bit_per_float: float = 32.0):
        # 0 < bit_per_float < 32: snapshots stored ZFP-coded at that rate (fixed-rate mode)
        z = (C.c_int * 3)(*nzones)
        h = lib().tti_sim3_create_ex(npoints, npml, niter, xcorr_step, z, int(fwd_only), float(bit_per_float))
        if not h:
            raise TTIError(f"simulator3SetUp: {lib().tti_host_last_error().decode()}")
        self._h = C.c_void_p(h)

    def close(self) -> None:
        if self._h:
            lib().tti_sim3_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_coefficients(self, df=None, si=None) -> None:
        """Stencil tables (4 values each; None keeps one). Zero by default, as the reference runs."""
        f4 = lambda v: None if v is None else (C.c_float * 4)(*[float(x) for x in v])
        check(lib().tti_sim3_set_coefficients(self._h, f4(df), f4(si)), "set_coefficients")

    def set_free_surface(self, on: bool) -> None:
        check(lib().tti_sim3_set_free_surface(self._h, int(on)), "set_free_surface")

    def set_snapshot_store(self, scratch: str = ".", max_mem_steps: int = -1) -> None:
        """Snapshot steps beyond max_mem_steps (or the host-RAM budget) go to a scratch file; before forward()."""
        check(lib().tti_sim3_set_snapshot_store(self._h, scratch.encode(), max_mem_steps), "set_snapshot_store")

    def execute(self, has_src_p=True, has_qp=False) -> None:
        check(lib().tti_sim3_execute(self._h, int(has_src_p), int(has_qp)), "execute")

    def forward(self, has_src_p=True) -> None:
        check(lib().tti_sim3_forward(self._h, int(has_src_p)), "forward")

    def adjoint(self) -> None:
        check(lib().tti_sim3_adjoint(self._h), "adjoint")

    @property
    def num_subs(self) -> int:
        return lib().tti_sim3_num_subs(self._h)

    def info(self, name: str, isub: int) -> dict:
        a = (C.c_int * 12)()
        check(lib().tti_sim3_volume_info(self._h, name.encode(), isub, a), "volume_info")
        v = list(a)
        return {"offset": v[0:3], "n": v[3:6], "ntot": v[6:9], "ibeg": v[9:12]}

    def read(self, name: str, isub: int) -> np.ndarray:
        """Whole allocation of a volume, shape (ntot_z, ntot_y, ntot_x); zprime: (ntot_z,)."""
        if name == "zprime":
            out = np.empty(self.info("f1", isub)["ntot"][2], dtype=np.float32)
        else:
            nt = self.info(name, isub)["ntot"]
            out = np.empty((nt[2], nt[1], nt[0]), dtype=np.float32)
        check(lib().tti_sim3_read(self._h, name.encode(), isub, out.ctypes.data), "read")
        return out

    def timings(self) -> dict:
        t = (C.c_double * 6)()
        check(lib().tti_sim3_timings(self._h, t), "timings")
        return dict(zip(("forward", "adjoint", "velocity", "strain", "stress", "halo"), list(t)))


class SnapStore:
    """Last-in first-out snapshot store (file_snap_orig): steps in host RAM up to mem_perc of the host memory, the rest
    in a temporary scratch file (POSIX AIO, O_DIRECT where allowed). max_mem_steps caps the steps kept in RAM."""

    def __init__(self, nval: int, nt: int, mem_perc: float = 0.88, scratch: str | None = None,
                 max_mem_steps: int = -1):
        if scratch is None and max_mem_steps < 0:
            h = lib().tti_snap_store_create(nval, mem_perc, nt)
        else:
            h = lib().tti_snap_store_create_ex(nval, mem_perc, nt, (scratch or ".").encode(), max_mem_steps)
        if not h:
            raise TTIError(f"file_snap: {lib().tti_host_last_error().decode()}")
        self._h = C.c_void_p(h)
        self.nval = nval
        self.nt = nt

    def strategy(self) -> tuple[str, bool]:
        """('.' memory / '#' disk per step, whether the scratch file uses O_DIRECT)."""
        buf = C.create_string_buffer(self.nt + 1)
        direct = C.c_int(0)
        check(lib().tti_snap_store_strategy(self._h, buf, self.nt + 1, C.byref(direct)), "strategy")
        return buf.value.decode(), bool(direct.value)

    def write(self, step: np.ndarray) -> None:
        a = np.ascontiguousarray(step, dtype=np.float32)
        if a.size != self.nval:
            raise ValueError("step size differs from the store's")
        check(lib().tti_snap_store_write(self._h, a.ctypes.data), "write_step")

    def read(self) -> np.ndarray:
        out = np.empty(self.nval, dtype=np.float32)
        check(lib().tti_snap_store_read(self._h, out.ctypes.data), "read_step")
        return out

    def close(self) -> None:
        if self._h:
            lib().tti_snap_store_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
