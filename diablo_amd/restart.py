"""Restart files with the datasets of the reference's WriteHDF5 / ReadHDF5
(for/hdf5.f:5-544), stored as a NumPy .npz archive (no HDF5 library is
needed): "Resolution" (int32 NX, NY, NZ; for/hdf5.f:112-118), "Time"
(for/hdf5.f:161-166), and the physical FP64 fields "U", "V", "W", "TH1".., and
"P" when saved.  Each field array has shape (NZ, NY, NX) in C order -- the
bytes and the shape an HDF5 reader shows for the reference's Fortran
(NX, NY, NZ) datasets -- with V on the fractional grid, as WriteHDF5 writes it.
The conversions run on the device (dgpu_get_flow_physical /
dgpu_set_flow_physical)."""
from __future__ import annotations

import numpy as np

from .gpu import CP, CTH0, CU1, CU2, CU3, GpuChannel

_VEL = (("U", CU1), ("V", CU2), ("W", CU3))


def save_flow(path: str, chan: GpuChannel, save_pressure: bool = True) -> None:
    """WriteHDF5's datasets of the device state (which is not changed)."""
    cfg = chan.cfg
    out = {"Resolution": np.array([cfg.nx, cfg.ny, cfg.nz], dtype=np.int32),
           "Time": np.array(chan.time, dtype=np.float64)}
    for name, fid in _VEL:
        out[name] = chan.get_flow_physical(fid)
    for n in range(cfg.n_th):
        out[f"TH{n + 1}"] = chan.get_flow_physical(CTH0 + n)
    if save_pressure:
        out["P"] = chan.get_flow_physical(CP)
    np.savez(path, **out)


def load_flow(path: str, chan: GpuChannel, create_new_th: tuple[bool, ...] = ()) -> None:
    """READ_FLOW with ReadHDF5 (for/diablo_io.f:201-301, for/hdf5.f:380-544):
    the Resolution check, Time, the fields (a scalar n is read unless
    create_new_th[n] is set, as CREATE_NEW_TH), POISSON_P_CHAN when the file
    holds no P, then the velocity BCs (APPLY_BC_VEL), which set the ghost
    planes the file does not hold.  As in the reference, the scalars' ghost
    planes stay zero until the first scalar solve."""
    cfg = chan.cfg
    with np.load(path, allow_pickle=False) as f:
        res = tuple(int(v) for v in f["Resolution"])
        if res != (cfg.nx, cfg.ny, cfg.nz):
            raise ValueError(f"file and program have different resolutions: file {res}, "
                             f"program {(cfg.nx, cfg.ny, cfg.nz)}")
        chan.time = float(f["Time"])
        for name, fid in _VEL:
            chan.set_flow_physical(fid, f[name])
        for n in range(cfg.n_th):
            if not (n < len(create_new_th) and create_new_th[n]):
                chan.set_flow_physical(CTH0 + n, f[f"TH{n + 1}"])
        has_p = "P" in f.files
        if has_p:
            chan.set_flow_physical(CP, f["P"])
    if not has_p:
        chan.poisson_p_chan()
    chan.apply_bc_vel()
