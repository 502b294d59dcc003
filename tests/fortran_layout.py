"""Tables of an asr_amd operator written in the reference's own layout (the fields of augcls_op, src/asr_augcls.f90:45-68,
and the blocks of class_alloy_t: 1-based nodes, operations, orbitals, classes, types and species, 0 = no link, column-major
arrays) for the Fortran caller asr_amd/fortran/asr_gpu_demo, which translates them with gpu_op_from_tables."""
import numpy as np

from asr_amd import engine


def _w(f, a, fmt="%d"):
    a = np.asarray(a).reshape(-1)
    for k in range(0, a.size, 12):
        f.write(" ".join(fmt % v for v in a[k:k + 12]) + "\n")


def write_reference_layout(path, op, members, nlev):
    with open(path, "w") as f:
        f.write("%d %d %d %d %d %d %d %d\n" % (op.nn, op.nslot, op.nop, op.nsmax, op.ncls, op.nsp, op.ntype, op.stab.size))
        has_h = op.hopr >= 0
        has_f = op.flipr >= 0
        _w(f, np.where(has_h, op.hopr + 1, 0))          # hopr(slot, node): node fastest index is the slot
        _w(f, np.where(has_h, op.hopg + 1, 1))
        _w(f, np.where(has_f, op.flipr + 1, 0))         # flipr(state + 1, node)
        _w(f, np.where(has_f, op.flipg + 1, 1))
        _w(f, op.lst)                                    # 0-based local state, as augcls_op%lst
        _w(f, op.ncl + 1)
        _w(f, op.ntyp + 1)                               # stype(site(i))
        _w(f, (op.fz == 1).astype(int))
        _w(f, op.sq, "%.17e")
        _w(f, op.stab_ptr + 1)
        _w(f, op.stab + 1)
        _w(f, op.pi + 1)                                 # pi(l, g)
        _w(f, op.sg, "%.17e")
        _w(f, np.transpose(op.block, (0, 1, 3, 2)), "%.17e")   # block(l, l2, slot, type), l fastest
        _w(f, np.transpose(op.block0, (0, 2, 1)), "%.17e")
        _w(f, op.nsc)
        _w(f, np.where(op.spc >= 0, op.spc + 1, 0))      # spc(k, class)
        _w(f, np.transpose(op.mt, (0, 3, 2, 1)), "%.17e")      # mt(l + 1, k + 1, s, class), l fastest
        s, p, fz, _, pab, z = engine.pack_members(op, members, complex_route=True)
        f.write("%d %d\n" % (len(members), nlev))
        _w(f, s)
        _w(f, fz)
        _w(f, p, "%.17e")
        _w(f, pab, "%.17e")
        _w(f, np.stack([z.real, z.imag], axis=1), "%.17e")


def read_results(path, nchain, nlev):
    mu = np.zeros(nchain, int)
    mr = np.zeros(nchain, int)
    a = np.zeros((nchain, nlev), complex)
    b = np.zeros((nchain, nlev), complex)
    ar = np.zeros((nchain, nlev))
    br = np.zeros((nchain, nlev))
    with open(path) as f:
        for j in range(nchain):
            mu[j], mr[j] = map(int, f.readline().split())
            for k in range(nlev):
                v = [float(x) for x in f.readline().split()]
                a[j, k] = complex(v[0], v[1])
                b[j, k] = complex(v[2], v[3])
                ar[j, k], br[j, k] = v[4], v[5]
    return mu, a, b, mr, ar, br
