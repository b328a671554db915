"""Isotropic linear elastic material (fealpy/material/elastic_material.py:148-319).

Accepts (elastic_modulus, poisson_ratio) or (lame_lambda, shear_modulus), or
all four when they agree; ``elastic_matrix()`` returns D of shape (1, 1, 6, 6)
('3D') or (1, 1, 3, 3) ('plane_strain', 'plane_stress') in FP64.
"""
import torch


class LinearElasticMaterial:
    def __init__(self, name, elastic_modulus=None, poisson_ratio=None, lame_lambda=None, shear_modulus=None,
                 density=None, hypo="3D", device=None):
        self.name, self.hypo, self.device = name, hypo, device
        E, nu, lam, mu = elastic_modulus, poisson_ratio, lame_lambda, shear_modulus
        if E is not None and nu is not None and lam is None and mu is None:
            lam = nu * E / ((1 + nu) * (1 - 2 * nu))
            mu = E / (2 * (1 + nu))
        elif lam is not None and mu is not None and E is None and nu is None:
            E = mu * (3 * lam + 2 * mu) / (lam + mu)
            nu = lam / (2 * (lam + mu))
        elif None not in (E, nu, lam, mu):
            cE = mu * (3 * lam + 2 * mu) / (lam + mu)
            cnu = lam / (2 * (lam + mu))
            if abs(cE - E) > 1e-5 or abs(cnu - nu) > 1e-5:
                raise ValueError("The input elastic modulus and Poisson's ratio are inconsistent with "
                                 "the values calculated from the provided Lame's lambda and shear modulus.")
        else:
            raise ValueError("You must provide either (elastic_modulus, poisson_ratio) or "
                             "(lame_lambda, shear_modulus), or all four.")
        self.E, self.nu, self.lam, self.mu, self.rho = float(E), float(nu), float(lam), float(mu), density
        if hypo == "3D":
            D = [[2 * mu + lam, lam, lam, 0, 0, 0], [lam, 2 * mu + lam, lam, 0, 0, 0],
                 [lam, lam, 2 * mu + lam, 0, 0, 0], [0, 0, 0, mu, 0, 0], [0, 0, 0, 0, mu, 0], [0, 0, 0, 0, 0, mu]]
            self.D = torch.tensor(D, dtype=torch.float64)
        elif hypo == "plane_stress":
            self.D = E / (1 - nu ** 2) * torch.tensor([[1, nu, 0], [nu, 1, 0], [0, 0, (1 - nu) / 2]],
                                                      dtype=torch.float64)
        elif hypo == "plane_strain":
            self.D = torch.tensor([[2 * mu + lam, lam, 0], [lam, 2 * mu + lam, 0], [0, 0, mu]], dtype=torch.float64)
        else:
            raise NotImplementedError("Only 3D, plane_stress, and plane_strain are supported.")

    @property
    def lame_lambda(self):
        return self.lam

    @property
    def shear_modulus(self):
        return self.mu

    @property
    def elastic_modulus(self):
        return self.E

    @property
    def poisson_ratio(self):
        return self.nu

    def elastic_matrix(self, bcs=None):
        return self.D[None, None, ...]
