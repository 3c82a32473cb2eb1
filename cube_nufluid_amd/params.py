"""Run-time form of CUBE's compile-time parameters (main/parameters.f90).

One image (nn=1): the simulation box is the image.  Sizes follow
parameters.f90:24-54; cosmology :70-107; particle compression :8-19,111-116.
"""
import dataclasses
import os

import numpy as np

NCELL = 4          # parameters.f90:27
NF_CUTOFF = 16     # :52
PI_F = np.float32(3.14159274101257324)  # pi=4*atan(1.) in single precision (:67)
DATA_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")


@dataclasses.dataclass
class BoxConfig:
    """Geometry, compression and cosmology of one CUBE box."""
    nc: int                 # coarse cells per dim (:29)
    nnt: int                # tiles per dim (:28)
    ncb: int = 6            # coarse buffer cells (:47)
    izipx: int = 2          # (:8)
    izipv: int = 2          # (:9)
    np_nc: int = 2          # particles per coarse cell per dim
    box: float = 100.0      # Mpc/h (:70)
    pp_on: bool = False     # Extended_pp_force (:60)
    pp_range: int = 1       # (:62)
    rsoft: float = 0.3      # (:61)
    h0: float = 0.68
    omega_m: float = 0.31
    omega_b: float = 0.049  # shape of the Eisenstein-Hu no-wiggle transfer function only
    n_s: float = 0.96
    sigma8: float = 0.8
    wde: float = -1.0       # (:100)
    z_i: float = 49.0
    np_max_factor: float = 1.0
    # neutrino fluid (parameters.f90:73-94): mass of each of the dneu = 3
    # degenerate species in eV; 0 runs CDM alone
    m_nu_ev: float = 0.0
    z_i_nu: float = 10.0

    @property
    def nt(self):
        return self.nc // self.nnt

    @property
    def nf(self):
        return self.nc * NCELL

    @property
    def omega_l(self):
        # parameters.f90:97-99 with zeq=1e9 (omega_r ~ 0)
        omega_r = self.omega_m / (1.0 + 1e9)
        return 1.0 - self.omega_m - omega_r

    @property
    def f_neu(self):
        """omega_nu/omega_m with omega_nu = dneu*Mneu/93.14/h0^2, dneu = 3
        (parameters.f90:81-93); omega_m is the total matter density."""
        return 3.0 * self.m_nu_ev / 93.14 / self.h0 ** 2 / self.omega_m

    @property
    def f_cdm(self):
        """omega_cdm/omega_m (parameters.f90:94): the matter that is not in
        the neutrino fluid (no gas)."""
        return 1.0 - self.f_neu

    @property
    def npart(self):
        return (self.nc * self.np_nc) ** 3

    @property
    def np_max(self):
        return int(self.npart * self.np_max_factor)

    def mass_p(self, npglobal=None):
        # util/initial_conditions.f90:943: f_cdm*nf_global**3/npglobal (no neutrinos: f_cdm=1)
        n = self.npart if npglobal is None else npglobal
        return np.float32(np.float32(self.f_cdm) * np.float32(float(self.nf) ** 3) / np.float32(n))


# Workloads of the benchmark configurations (seeded, EH no-wiggle LCDM).
CONFIGS = {
    "C1": BoxConfig(nc=32, nnt=2, ncb=6, box=100.0, pp_on=False),
    "C1pp": BoxConfig(nc=32, nnt=2, ncb=6, box=100.0, pp_on=True, pp_range=1),
    "C2": BoxConfig(nc=128, nnt=4, ncb=6, izipx=1, izipv=2, box=400.0, pp_on=True, pp_range=1),
    "C3": BoxConfig(nc=512, nnt=4, ncb=6, izipx=2, izipv=2, box=1000.0, pp_on=True, pp_range=1),
    # CDM + neutrino fluid: sum m_nu = 0.1 eV as three degenerate species,
    # fluids on 512^3, 256^3 and 128^3 cells, started at z = 10
    "C4": BoxConfig(nc=256, nnt=4, ncb=6, izipx=2, izipv=2, box=500.0, pp_on=True, pp_range=1, z_i=10.0,
                    m_nu_ev=0.1 / 3, z_i_nu=10.0),
}


def read_kernel_tables(data_dir=DATA_DIR):
    """kernels/wfxyzf.3.ascii (kernel_f.f90:14-23) and wfxyzc.2.ascii
    (kernel_c.f90:31-39) as float32 arrays of 4096x3 and 64x3 in file order."""
    kf = np.loadtxt(os.path.join(data_dir, "wfxyzf.3.ascii"), dtype=np.float64)
    kc = np.loadtxt(os.path.join(data_dir, "wfxyzc.2.ascii"), dtype=np.float64)
    if kf.shape != (4096, 6) or kc.shape != (64, 6):
        raise ValueError("unexpected kernel table shapes %s %s" % (kf.shape, kc.shape))
    # index check of kernel_f.f90:19 (rows are i fastest, then j, then k)
    idx = np.arange(4096)
    if not (np.all(kf[:, 0] == idx % 16 + 1) and np.all(kf[:, 1] == (idx // 16) % 16 + 1)
            and np.all(kf[:, 2] == idx // 256 + 1)):
        raise ValueError("wfxyzf.3.ascii index columns out of order")
    return (np.ascontiguousarray(kf[:, 3:6], dtype=np.float32),
            np.ascontiguousarray(kc[:, 3:6], dtype=np.float32))
