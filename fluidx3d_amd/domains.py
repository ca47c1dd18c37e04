"""A lattice split into Dx*Dy*Dz domains with one-cell halos, one C-ABI context per domain:
the multi-domain LBM of the reference (src/lbm.cpp:721-773 construction, 881-953
initialize/do_time_step, 1355-1383 communicate_field) on the halo-transfer entry points.

    lbm = DomainLBM(64, 48, 40, 2, 1, 2, nu=0.05)   # LBM(Nx,Ny,Nz,Dx,Dy,Dz,nu,...)
    lbm.rho[:] = ...; lbm.u[:] = ...; lbm.flags[:] = ...   # global host arrays
    lbm.run(100)
    rho, u = lbm.read_fields()                              # interior cells, global layout

Every domain is a box of (N/D + 2) cells along each split axis (src/lbm.cpp:727-733) with
its own DDFs, rho/u/flags and stream.  After each stream_collide the DDFs that cross a
face are extracted into per-domain buffers and inserted by the neighbour, axis by axis.
All domains here live on one device and read each other's buffers directly; on several
devices the same buffers would move device to device (see the sharding notes of the
design).  The results on the interior cells are those of the undivided lattice, bit for
bit.
"""
import numpy as np

from . import _abi
from ._abi import check
from .lbm import LBM


class DomainLBM:
    def __init__(self, Nx, Ny, Nz, Dx, Dy, Dz, nu, fx=0.0, fy=0.0, fz=0.0, *, velocity_set=19, collision="SRT",
                 storage="FP32", extensions=(), device=0):
        D = (int(Dx), int(Dy), int(Dz))
        if min(D) < 1:
            raise ValueError("domain counts must be at least 1")
        N3 = [(int(n) // d) * d for n, d in zip((Nx, Ny, Nz), D)]  # divisible by the domains (src/lbm.cpp:724-726)
        self.Nx, self.Ny, self.Nz = N3
        self.N = self.Nx * self.Ny * self.Nz
        self.D = D
        self.H = tuple(1 if d > 1 else 0 for d in D)
        self.local = tuple(n // d + 2 * h for n, d, h in zip(N3, D, self.H))
        if any(d > 1 and n // d < 1 for n, d in zip(N3, D)):
            raise ValueError("every domain needs at least one interior cell along a split axis")
        self.velocity_set = int(velocity_set)
        self.storage = storage
        self.fx, self.fy, self.fz = float(fx), float(fy), float(fz)
        w = _abi.relaxation_rate(nu)
        self.domains = []
        self.offsets = []
        for dz in range(D[2]):
            for dy in range(D[1]):
                for dx in range(D[0]):  # d = x + (y + z*Dy)*Dx
                    self.offsets.append((dx * self.Nx // D[0] - self.H[0], dy * self.Ny // D[1] - self.H[1],
                                         dz * self.Nz // D[2] - self.H[2]))
                    self.domains.append(LBM(*self.local, nu, fx, fy, fz, velocity_set=velocity_set, collision=collision,
                                            storage=storage, extensions=extensions, device=device, w=w))
        self.has_F = self.domains[0].F is not None
        # global host fields in the reference layout
        self.rho = np.ones(self.N, np.float32)
        self.u = np.zeros(3 * self.N, np.float32)
        self.flags = np.zeros(self.N, np.uint8)
        self.F = np.zeros(3 * self.N, np.float32) if self.has_F else None
        # per domain and split axis: (buffer_p, buffer_m), sized for the largest field
        self.buffers = []
        for L in self.domains:
            per_axis = {}
            for axis in range(3):
                if D[axis] > 1:
                    nbytes = max(int(L.ctx.lib.fx_transfer_bytes(L.cfg, f, axis)) for f in range(4))
                    per_axis[axis] = (L.ctx.alloc(nbytes), L.ctx.alloc(nbytes))
            self.buffers.append(per_axis)
        self.t = 0
        self.initialized = False
        # global cell index of every local cell of a domain (periodic wrap of the halos)
        lx, ly, lz = self.local
        self._gidx = []
        self._interior = []
        for (ox, oy, oz) in self.offsets:
            z, y, x = np.meshgrid(np.arange(lz), np.arange(ly), np.arange(lx), indexing="ij")
            gx, gy, gz = (x + ox) % self.Nx, (y + oy) % self.Ny, (z + oz) % self.Nz
            self._gidx.append((gx + (gy + gz * self.Ny) * self.Nx).ravel())
            inner = np.ones(x.shape, bool)
            for axis, c, n in ((0, x, lx), (1, y, ly), (2, z, lz)):
                if self.H[axis]:
                    inner &= (c >= 1) & (c <= n - 2)
            self._interior.append(inner.ravel())

    def neighbour(self, d, axis, step):
        Dx, Dy, Dz = self.D
        c = [d % Dx, (d // Dx) % Dy, d // (Dx * Dy)]
        c[axis] = (c[axis] + step) % self.D[axis]
        return c[0] + (c[1] + c[2] * Dy) * Dx

    # -- communication (src/lbm.cpp:1355-1383) -----------------------------------------
    def _sync(self):
        for L in self.domains:
            L.ctx.sync()

    def _communicate(self, field):
        for axis in range(3):
            if self.D[axis] == 1:
                continue
            for d, L in enumerate(self.domains):
                bp, bm = self.buffers[d][axis]
                check(L.ctx.lib.fx_lbm_transfer_extract(L.ctx.h, field, axis, self.t, bp, bm, L.fi, L.rho.device, L.u.device,
                                                         L.flags.device, L._ptr(L.F)))
            self._sync()  # every extract done before any insert reads its buffers
            for d, L in enumerate(self.domains):
                from_p = self.buffers[self.neighbour(d, axis, +1)][axis][1]  # + neighbour's side -
                from_m = self.buffers[self.neighbour(d, axis, -1)][axis][0]  # - neighbour's side +
                check(L.ctx.lib.fx_lbm_transfer_insert(L.ctx.h, field, axis, self.t, from_p, from_m, L.fi, L.rho.device, L.u.device,
                                                        L.flags.device, L._ptr(L.F)))
            self._sync()  # inserts done before the buffers are written again

    # -- simulation ---------------------------------------------------------------------
    def initialize(self):
        """Interior fields to the devices, halos by transfer, initialize with t=1 on every
        domain, halo DDFs, t=0 (src/lbm.cpp:881-922)."""
        for d, L in enumerate(self.domains):
            g, inner = self._gidx[d], self._interior[d]
            Nl = L.N
            L.rho.host[:] = 1.0
            L.u.host[:] = 0.0
            L.flags.host[:] = 0
            L.rho.host[inner] = self.rho[g[inner]]
            for c in range(3):
                L.u.host[c * Nl:(c + 1) * Nl][inner] = self.u[c * self.N + g[inner]]
                if self.has_F:
                    L.F.host[c * Nl:(c + 1) * Nl] = self.F[c * self.N + g]
            L.flags.host[inner] = self.flags[g[inner]]
            L.rho.write_to_device()
            L.u.write_to_device()
            L.flags.write_to_device()
            if self.has_F:
                L.F.write_to_device()
        self.t = 1  # the communication at initialization needs an odd time step
        self._communicate(_abi.FX_FIELD_RHO_U_FLAGS)
        for L in self.domains:
            check(L.ctx.lib.fx_lbm_initialize(L.ctx.h, L.fi, L.rho.device, L.u.device, L.flags.device))
        self._communicate(_abi.FX_FIELD_RHO_U_FLAGS)
        self._communicate(_abi.FX_FIELD_FI)
        self._sync()
        self.t = 0
        for L in self.domains:
            L.initialized = True
            L.t = 0
        self.initialized = True

    def run(self, steps=1):
        """do_time_step (src/lbm.cpp:924-953) `steps` times."""
        if not self.initialized:
            self.initialize()
        for _ in range(int(steps)):
            for L in self.domains:
                L.t = self.t
                check(L.ctx.lib.fx_lbm_stream_collide(L.ctx.h, L.fi, L.rho.device, L.u.device, L.flags.device, self.t,
                                                       self.fx, self.fy, self.fz, L._ptr(L.F), L._ptr(L.porosity)))
            self._communicate(_abi.FX_FIELD_FI)
            self.t += 1
        for L in self.domains:
            L.t = self.t

    def read_fields(self):
        """update_fields on every domain; rho (N) and u (3N) of the interior cells."""
        rho = np.empty(self.N, np.float32)
        u = np.empty(3 * self.N, np.float32)
        for d, L in enumerate(self.domains):
            L.t = self.t
            L.rho.read_from_device()
            L.u.read_from_device()
            g, inner = self._gidx[d], self._interior[d]
            rho[g[inner]] = L.rho.host[inner]
            for c in range(3):
                u[c * self.N + g[inner]] = L.u.host[c * L.N:(c + 1) * L.N][inner]
        return rho, u

    def global_ddfs(self):
        """The DDF array of the undivided lattice (slot layout fi[i*N+n]) from the interior
        cells of the domains."""
        out = None
        for d, L in enumerate(self.domains):
            local = L.read_ddfs()
            if out is None:
                out = np.empty(self.velocity_set * self.N, local.dtype)
            g, inner = self._gidx[d], self._interior[d]
            for i in range(self.velocity_set):
                out[i * self.N + g[inner]] = local[i * L.N:(i + 1) * L.N][inner]
        return out

    def close(self):
        for L in self.domains:
            L.close()
