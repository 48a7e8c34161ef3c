"""Drop-in for LIBRARIES/gpu_Verlet.py: the same Verlet_GPU signature
(gpu_Verlet.py:53-61), the same in-place semantics on the caller's arrays and
the same return order (gpu_Verlet.py:89-92). The integrator is the HIP kernel
behind mwd_verlet (include/mwdiem_hip.h); nothing is compiled per call.

Arrays may be numpy (as the unchanged driver passes them; they are uploaded,
integrated and written back in place) or FP64 CUDA tensors (updated in place
on the device).
"""
from __future__ import annotations

import ctypes as C

import torch

from .. import _lib
from ._common import device, stream_ptr, to_dev, write_back

_F64 = torch.float64


def Verlet_GPU(velocity_arr, velocity_12_arr, angular_velocity_12_bf_arr,
               quaternion_12_arr, quaternion_arr, quaternion_all_arr,
               angular_velocity_bf_arr, angular_velocity_arr, force_arr,
               torque_arr, vertices_arr, center_of_mass_arr,
               timestep_size, mass_arr, num_particles,
               moment_of_inertia_inv_arr, moment_of_inertia_arr, dx_all_arr,
               num_vertices_max, syssize_x_max, syssize_x_min,
               syssize_y_max, syssize_y_min, syssize_z_max,
               syssize_z_min, num_part_min, num_part_max):
    dev = device()
    io = [velocity_arr, velocity_12_arr, angular_velocity_12_bf_arr, quaternion_12_arr, quaternion_arr,
          quaternion_all_arr, angular_velocity_bf_arr, angular_velocity_arr]
    t_io = [to_dev(a, _F64, dev) for a in io]
    F = to_dev(force_arr, _F64, dev)
    T = to_dev(torque_arr, _F64, dev)
    verts = to_dev(vertices_arr, _F64, dev)
    com = to_dev(center_of_mass_arr, _F64, dev)
    mass = to_dev(mass_arr, _F64, dev).reshape(-1)
    Iinv = to_dev(moment_of_inertia_inv_arr, _F64, dev)
    I = to_dev(moment_of_inertia_arr, _F64, dev)
    dxall = to_dev(dx_all_arr, _F64, dev)
    n_rows = com.shape[0]
    vmax = int(num_vertices_max)
    if verts.numel() != n_rows * vmax * 3:
        raise ValueError("vertices_arr must be [N, num_vertices_max, 3]")
    pmin = max(0, int(num_part_min))
    pmax = min(n_rows, int(num_part_max))
    box = (C.c_double * 6)(syssize_x_max, syssize_x_min, syssize_y_max, syssize_y_min, syssize_z_max,
                           syssize_z_min)
    v, v12, w12bf, q12, q, qall, wbf, w = t_io
    rc = _lib.load().mwd_verlet(v.data_ptr(), v12.data_ptr(), w12bf.data_ptr(), q12.data_ptr(), q.data_ptr(),
                                qall.data_ptr(), wbf.data_ptr(), w.data_ptr(), F.data_ptr(), T.data_ptr(),
                                verts.data_ptr(), com.data_ptr(), mass.data_ptr(), Iinv.data_ptr(), I.data_ptr(),
                                dxall.data_ptr(), vmax, None, float(timestep_size), box, pmin, pmax, stream_ptr())
    _lib.check(rc, "mwd_verlet")
    for dst, src in zip(io, t_io):
        write_back(dst, src)
    write_back(vertices_arr, verts)
    write_back(center_of_mass_arr, com)
    write_back(dx_all_arr, dxall)
    return (velocity_12_arr, quaternion_arr, angular_velocity_bf_arr,
            angular_velocity_12_bf_arr, angular_velocity_arr, quaternion_12_arr,
            velocity_arr, center_of_mass_arr, vertices_arr,
            quaternion_all_arr, dx_all_arr)
