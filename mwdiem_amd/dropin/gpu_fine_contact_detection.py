"""Drop-in for LIBRARIES/gpu_fine_contact_detection.py: the same
Contact_detection_GPU signature (gpu_fine_contact_detection.py:76-89) and
return value (force, torque, displacement_tang, collision) (:119).

GJK, EPA and the contact law run in FP64 in the HIP kernels behind
mwd_contact_pairs (include/mwdiem_hip.h) on the caller's per-pair arrays;
nothing is compiled per call. As with the reference's cuda.InOut arguments,
`force`, `torque` and `displacement_tang` are updated in place (and
returned): per particle, the contact sum in ascending pair order is added to
the incoming rows. `collision` is a new int32 array (-1 undecided, 0 apart,
1 contact). displacement_tang rows are [x, y, z, 1] for contacts and zeros
otherwise, stored at [4P+k] (the reference kernel's [P+k] indexing overlaps
neighbouring pairs; the driver discards these values every step,
MWDiem.py:966).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from .. import _lib
from ._common import device, stream_ptr, to_dev, write_back

_F64 = torch.float64
_I32 = torch.int32


def Contact_detection_GPU(particle_A_center_of_mass, particle_B_center_of_mass,
                          particle_A_vertices, particle_B_vertices,
                          particle_A_velocity, particle_B_velocity,
                          particle_A_angular_velocity, particle_B_angular_velocity,
                          particle_A_numvertices, particle_B_numvertices,
                          particle_A_id, particle_B_id,
                          displacement_tang,
                          Kn, en, et, mu,
                          Kn_wall, en_wall, et_wall, mu_wall, num_particles_wall,
                          force, torque,
                          num_vertices_maxi,
                          quaternion_A, quaternion_B,
                          timestep_size,
                          mom_inertia_size_A, mom_inertia_size_B):
    dev = device()
    P = int(np.asarray(particle_A_id).shape[0]) if not isinstance(particle_A_id, torch.Tensor) \
        else int(particle_A_id.shape[0])
    vmax = int(num_vertices_maxi)
    comA = to_dev(particle_A_center_of_mass, _F64, dev)
    comB = to_dev(particle_B_center_of_mass, _F64, dev)
    vA = to_dev(particle_A_vertices, _F64, dev)
    vB = to_dev(particle_B_vertices, _F64, dev)
    if vA.numel() != P * vmax * 3 or vB.numel() != P * vmax * 3:
        raise ValueError("particle vertices must be [P, num_vertices_maxi, 3]")
    velA = to_dev(particle_A_velocity, _F64, dev)
    velB = to_dev(particle_B_velocity, _F64, dev)
    wA = to_dev(particle_A_angular_velocity, _F64, dev)
    wB = to_dev(particle_B_angular_velocity, _F64, dev)
    nA = to_dev(particle_A_numvertices, _I32, dev)
    nB = to_dev(particle_B_numvertices, _I32, dev)
    idA = to_dev(particle_A_id, _I32, dev)
    idB = to_dev(particle_B_id, _I32, dev)
    disp = to_dev(displacement_tang, _F64, dev).reshape(P, 4) if P else torch.zeros((0, 4), dtype=_F64, device=dev)
    qA = to_dev(quaternion_A, _F64, dev)
    isA = to_dev(mom_inertia_size_A, _F64, dev)
    isB = to_dev(mom_inertia_size_B, _F64, dev)
    F = to_dev(force, _F64, dev)
    T = to_dev(torque, _F64, dev)
    n_particles = F.shape[0]
    if P and (int(torch.minimum(idA.min(), idB.min())) < 0 or int(torch.maximum(idA.max(), idB.max())) >= n_particles):
        raise ValueError("particle ids must index the force/torque rows")
    collision = torch.full((P,), -1, dtype=_I32, device=dev)
    disp_out = torch.zeros((P, 4), dtype=_F64, device=dev)
    lp = _lib.law(Kn, en, et, mu, timestep_size)
    lw = _lib.law(Kn_wall, en_wall, et_wall, mu_wall, timestep_size)
    if P:
        rc = _lib.load().mwd_contact_pairs(
            P, vmax, comA.data_ptr(), comB.data_ptr(), vA.data_ptr(), vB.data_ptr(), velA.data_ptr(),
            velB.data_ptr(), wA.data_ptr(), wB.data_ptr(), nA.data_ptr(), nB.data_ptr(), idA.data_ptr(),
            idB.data_ptr(), disp.data_ptr(), qA.data_ptr(), isA.data_ptr(), isB.data_ptr(), C.byref(lp),
            C.byref(lw), int(num_particles_wall), collision.data_ptr(), None, None, None, None, None,
            disp_out.data_ptr(), F.data_ptr(), T.data_ptr(), n_particles, stream_ptr())
        _lib.check(rc, "mwd_contact_pairs")
    write_back(force, F)
    write_back(torque, T)
    if P:
        write_back(displacement_tang, disp_out)
    coll = collision if isinstance(particle_A_id, torch.Tensor) else collision.cpu().numpy()
    return (force, torque, displacement_tang, coll)
