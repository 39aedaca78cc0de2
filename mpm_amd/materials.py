"""Material factory keyed like the reference's registrations.

The reference registers ``LinearElastic2D/3D`` and ``MohrCoulomb2D/3D`` in a
``Factory<Material<Tdim>, unsigned, const Json&>`` (src/material.cc:31-53) and
reads their properties from the JSON ``materials`` array
(include/materials/infinitesimal_strain/linear_elastic.tcc:3-49,
mohr_coulomb.tcc:3-81). Here a material is a 16-double parameter vector that
crosses the C-ABI unchanged (17 doubles):

    [0] type (0 LinearElastic, 1 MohrCoulomb)  [1] density  [2] youngs_modulus
    [3] poisson_ratio  [4] friction (deg)  [5] dilation (deg)  [6] cohesion
    [7] residual_friction (deg)  [8] residual_dilation (deg)
    [9] residual_cohesion  [10] peak_pdstrain  [11] residual_pdstrain
    [12] tension_cutoff  [13] softening (0/1)  [14] packing_fraction_minimum
    [15] initial packing fraction (<= 0 means 1)  [16] stress rate

The objective stress rate (linear_elastic.tcc:8-17; MohrCoulomb inherits it,
infinitesimal_elasto_plastic.tcc:1-25): 0 none, 1 Jaumann, 2 Green-Naghdi,
in [4] for LinearElastic and in [16] for MohrCoulomb.
"""
import numpy as np

NPARAMS = 17
LINEAR_ELASTIC = 0
MOHR_COULOMB = 1

# Names of the MohrCoulomb state variables, in the reference's order
# (mohr_coulomb.tcc:116-118). yield_state: 0 elastic, 1 shear, 2 tensile,
# 3 separated.
MC_STATE_VARIABLES = ("yield_state", "phi", "psi", "cohesion", "tension_cutoff",
                      "epsilon", "rho", "theta", "pdstrain")
NSTATEVARS = len(MC_STATE_VARIABLES)


STRESS_RATES = {None: 0, "none": 0, "jaumann": 1, "green_naghdi": 2}


def linear_elastic(youngs_modulus, poisson_ratio, density, stress_rate=None):
    p = np.zeros(NPARAMS)
    p[0] = LINEAR_ELASTIC
    p[1] = density
    p[2] = youngs_modulus
    p[3] = poisson_ratio
    p[4] = STRESS_RATES[stress_rate]
    return p


def mohr_coulomb(youngs_modulus, poisson_ratio, density, friction, dilation,
                 cohesion, residual_friction=None, residual_dilation=None,
                 residual_cohesion=None, peak_pdstrain=0.0,
                 residual_pdstrain=0.0, tension_cutoff=0.0, softening=False,
                 packing_fraction=None, porosity=None,
                 packing_fraction_minimum=0.0, stress_rate=None):
    """MohrCoulomb parameters; residual values default to the peak ones.

    ``packing_fraction`` / ``porosity`` set the initial packing fraction as in
    mohr_coulomb.tcc:12-31 (single-phase: grain density = density / packing).
    """
    p = np.zeros(NPARAMS)
    p[0] = MOHR_COULOMB
    p[1] = density
    p[2] = youngs_modulus
    p[3] = poisson_ratio
    p[4] = friction
    p[5] = dilation
    p[6] = cohesion
    p[7] = friction if residual_friction is None else residual_friction
    p[8] = dilation if residual_dilation is None else residual_dilation
    p[9] = cohesion if residual_cohesion is None else residual_cohesion
    p[10] = peak_pdstrain
    p[11] = residual_pdstrain
    p[12] = tension_cutoff
    p[13] = 1.0 if softening else 0.0
    p[14] = packing_fraction_minimum
    if packing_fraction is not None:
        p[15] = packing_fraction
    elif porosity is not None:
        p[15] = 1.0 - porosity
    p[16] = STRESS_RATES[stress_rate]
    return p


_REQUIRED = {
    "LinearElastic": ("density", "youngs_modulus", "poisson_ratio"),
    "MohrCoulomb": ("density", "youngs_modulus", "poisson_ratio", "softening",
                    "friction", "dilation", "cohesion", "residual_friction",
                    "residual_dilation", "residual_cohesion", "peak_pdstrain",
                    "residual_pdstrain", "tension_cutoff"),
}


def from_json(props, dim):
    """Build a parameter vector from one entry of the JSON ``materials`` array.

    ``type`` must be one of the reference keys ``LinearElastic{2,3}D`` or
    ``MohrCoulomb{2,3}D`` for the analysis dimension; every property the
    reference reads with ``.at()`` is required here too.
    """
    mtype = props["type"]
    for base in _REQUIRED:
        if mtype == "%s%dD" % (base, dim):
            break
    else:
        raise ValueError("material type %r is not supported in %dD" % (mtype, dim))
    missing = [k for k in _REQUIRED[base] if k not in props]
    if missing:
        raise KeyError("Material parameter not set: %s" % ", ".join(missing))
    rate = props.get("stress_rate")
    rate = rate if rate in STRESS_RATES else None  # other names: no rate (linear_elastic.tcc:17)
    if base == "LinearElastic":
        return linear_elastic(props["youngs_modulus"], props["poisson_ratio"],
                              props["density"], stress_rate=rate)
    return mohr_coulomb(
        props["youngs_modulus"], props["poisson_ratio"], props["density"],
        props["friction"], props["dilation"], props["cohesion"],
        props["residual_friction"], props["residual_dilation"],
        props["residual_cohesion"], props["peak_pdstrain"],
        props["residual_pdstrain"], props["tension_cutoff"],
        bool(props["softening"]), props.get("packing_fraction"),
        props.get("porosity"), props.get("packing_fraction_minimum", 0.0), stress_rate=rate)
