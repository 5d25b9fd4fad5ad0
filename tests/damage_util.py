"""The two damage materials of the fixtures and tests (types 6 and 3): copper
J2 with isotropic damage, and a concrete-like Drucker-Prager cone with
damage, both with thresholds that the 190 m/s bar crosses within 40 steps."""
from taylor_amd import deck


def damage_material(mat_type):
    if mat_type == 6:
        return deck.Material.j2_damage(E=117e9, nu=0.35, rho=8930.0, sig0=400e6, B=0.25, CE=1.0,
                                       k_i=0.002, k_c=0.05, dam_max=0.8)
    return deck.Material.drucker_prager(E=30e9, nu=0.2, rho=2400.0, dp_type=1, phi_deg=30.0, c=10e6,
                                        psi_deg=0.0, t_cut=3e6, k_i=0.0005, k_c=0.01, dam_max=0.8)
