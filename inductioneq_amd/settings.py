"""Program configuration in the reference's vocabulary.

The reference assembles one OpenCL program per run from a list of source modules
(``matlab/+induction_eq/compute_numerical_solution_setup.m:57-122``) and a string of
``-D`` build options produced by ``generate_settings`` (``matlab/generate_settings.m:3-30``)
from the ``I_Mesh/I_TI/I_IEq/I_DC/I_Tech/I_RunOps`` maps
(``matlab/+induction_eq/prepare_vars.m:29-56``).  :class:`ProgramConfig` holds the same
settings as typed fields, renders the identical settings string and module list, and
parses them back; the C-ABI library parses the same string itself
(``ieq_program_build``).
"""
from __future__ import annotations

import dataclasses
import math
import re
from typing import Dict, List, Optional, Sequence

OPSET_MODULES = ["2ndOrder", "4thOrder", "6thOrder", "2ndOrderExtended", "4thOrderExtended", "6thOrderExtended"]

# reference testcases (include/<name>.h) plus two added for the benchmark workloads
TESTCASES = [
    "rotation_2D",            # include/rotation_2D.h
    "rotation_3D",            # include/rotation_3D.h
    "hall_periodic",          # include/hall_periodic.h
    "hall_travelling_wave",   # include/hall_travelling_wave.h
    "confined_domain",        # include/confined_domain.h
    "alfven_periodic_2D",     # include/alfven_periodic_2D.h
    "gaussian_rotation_3D",   # B0 = curl(0,0,exp(-20((x-1/2)^2+y^2+z^2))), u = (-y,x,0)
    "uniform_flow",           # host-initialised B, u = (UX,UY,UZ), b_boundary = 0
]

UIBJ_FORMS = ["CENTRAL", "SPLIT", "PRODUCT"]
SOURCE_FORMS = ["ZERO", "CENTRAL", "SPLIT"]
UJBI_FORMS = ["CENTRAL", "SPLIT", "PRODUCT"]
DISSIPATION_FORMS = {None: 0, "HIGH_ORDER": 1, "FIRST_ORDER": 2, "ADAPTIVE": 3}
LAPLACE_FORMS = {
    "WIDE_STENCIL_DIRICHLET": 0,  # USE_LAPLACE_WIDE_STENCIL_DIRICHLET (WS_D0)
    "NARROW_STENCIL_DIRICHLET": 1,  # USE_LAPLACE_NARROW_STENCIL_DIRICHLET (NS_D0)
    "WIDE_STENCIL_LNS": 2,  # USE_LAPLACE_WIDE_STENCIL_LNS (WS_LN)
}


def generate_settings(mapping: Dict[str, object], keys: Sequence[str]) -> str:
    """Render ``-D`` options like ``matlab/generate_settings.m:3-30``.

    Strings containing ``USE`` become ``-D<value>=1``; the key ``optimizations`` is
    appended verbatim; other strings become ``-D<key>=<value>``; integers use ``%d``
    and floating-point values ``%.16e``.  Unknown keys are reported and skipped.
    """
    out = ""
    for key in keys:
        if key not in mapping:
            print("Unknown identifier %s" % key)
            continue
        val = mapping[key]
        if isinstance(val, str):
            if key == "optimizations":
                out += val
            elif "USE" in val:
                out += " -D%s=1" % val
            else:
                out += " -D%s=%s" % (key, val)
        elif isinstance(val, bool):
            out += " -D%s=%d" % (key, int(val))
        elif isinstance(val, int):
            out += " -D%s=%d" % (key, val)
        else:
            out += " -D%s=%.16e" % (key, float(val))
    return out


@dataclasses.dataclass
class ProgramConfig:
    """All compile-time settings of one reference program, as run-time data."""

    real: str = "double"
    opset: str = "4thOrder"
    nx: int = 16
    ny: int = 16
    nz: int = 16
    xmin: float = -1.0
    xmax: float = 1.0
    ymin: float = -1.0
    ymax: float = 1.0
    zmin: float = -1.0
    zmax: float = 1.0
    dx: Optional[float] = None
    dy: Optional[float] = None
    dz: Optional[float] = None
    dt: float = 1e-3
    periodic: bool = False
    hall: bool = False
    dissipation: Optional[str] = None  # HIGH_ORDER | FIRST_ORDER | ADAPTIVE
    ho_factor: float = 1.0
    mp2min: float = 1.0
    mp2max: float = 10.0
    cmin: float = 1.0
    cmax: float = 10.0
    form_uibj: str = "PRODUCT"
    form_source: str = "CENTRAL"
    form_ujbi: str = "CENTRAL"
    testcase: str = "rotation_2D"
    laplace: Optional[str] = None  # LAPLACE_FORMS key; None = no divergence-cleaning module
    w_size: int = 256
    ux: float = 0.0
    uy: float = 0.0
    uz: float = 0.0

    def __post_init__(self):
        # grid spacing rule: matlab/+induction_eq/initialize.m:12-21
        if self.dx is None:
            self.dx = (self.xmax - self.xmin) / (self.nx if self.periodic else self.nx - 1)
        if self.dy is None:
            self.dy = (self.ymax - self.ymin) / (self.ny if self.periodic else self.ny - 1)
        if self.dz is None:
            self.dz = (self.zmax - self.zmin) / (self.nz if self.periodic else self.nz - 1)
        if self.opset not in OPSET_MODULES:
            raise ValueError("unknown operator module %r" % self.opset)
        if self.testcase not in TESTCASES:
            raise ValueError("unknown testcase %r" % self.testcase)

    # --- derived quantities -------------------------------------------------------
    @property
    def opset_index(self) -> int:
        return OPSET_MODULES.index(self.opset)

    @property
    def num_nodes(self) -> int:
        return self.nx * self.ny * self.nz

    @property
    def num_nodes_pad(self) -> int:
        # matlab/+induction_eq/initialize.m:43
        return int(math.ceil(self.num_nodes / self.w_size) * self.w_size)

    @property
    def num_groups(self) -> int:
        return self.num_nodes_pad // self.w_size

    @property
    def bnodes(self) -> int:
        # matlab/+induction_eq/initialize.m:107-114
        if self.laplace is None:
            return 0
        return 0 if self.laplace == "WIDE_STENCIL_LNS" else 1

    @property
    def laplace_index(self) -> int:
        return LAPLACE_FORMS[self.laplace] if self.laplace else 0

    @property
    def dtype(self):
        import numpy as np

        return np.float64 if self.real == "double" else np.float32

    # --- reference vocabulary -------------------------------------------------------
    def modules(self) -> List[str]:
        """The program's source list (compute_numerical_solution_setup.m:59-122)."""
        mods = [self.opset, "utils", self.testcase, "kernel_init", "SBP_operator", "kernel_operator"]
        if self.laplace is not None:
            mods += ["kernel_div_cleaning", "kernel_CG"]
        if self.dissipation is not None:
            mods += ["artificial_dissipation"]
        mods += ["induction_Eq_volume", "induction_Eq_surface", "kernel_dot_product", "kernel_time_integrator"]
        return mods

    def settings_string(self) -> str:
        """The ``-D`` build-option string (compute_numerical_solution_setup.m:125-137)."""
        ieq = {"form_uibj": "USE_UIBJ_" + self.form_uibj, "form_source": "USE_SOURCE_" + self.form_source,
               "form_ujbi": "USE_UJBI_" + self.form_ujbi, "hall_term": "USE_HALL" if self.hall else "NONE"}
        s = generate_settings(ieq, ["form_uibj", "form_source", "form_ujbi", "hall_term"])
        s += generate_settings({"DT": float(self.dt)}, ["DT"])
        tech = {"REAL": self.real, "REAL4": self.real + "4", "W_SIZE": int(self.w_size),
                "optimizations": " -cl-mad-enable -cl-no-signed-zeros -cl-finite-math-only"
                + (" -cl-single-precision-constant" if self.real == "float" else "")}
        s += generate_settings(tech, ["REAL", "REAL4", "W_SIZE", "optimizations"])
        mesh = {"DX": float(self.dx), "DY": float(self.dy), "DZ": float(self.dz), "NODES_X": int(self.nx),
                "NODES_Y": int(self.ny), "NODES_Z": int(self.nz), "XMIN": float(self.xmin), "XMAX": float(self.xmax),
                "YMIN": float(self.ymin), "YMAX": float(self.ymax), "ZMIN": float(self.zmin), "ZMAX": float(self.zmax)}
        s += generate_settings(mesh, ["DX", "DY", "DZ", "NODES_X", "NODES_Y", "NODES_Z", "XMIN", "XMAX", "YMIN", "YMAX",
                                      "ZMIN", "ZMAX"])
        if self.dissipation is not None:
            diss = {"dissipation": "USE_ARTIFICIAL_DISSIPATION", "dissipation_form": "USE_%s_DISSIPATION" % self.dissipation,
                    "HO_DISSIPATION_FACTOR": float(self.ho_factor)}
            s += generate_settings(diss, ["dissipation", "dissipation_form", "HO_DISSIPATION_FACTOR"])
            if self.dissipation == "ADAPTIVE":
                s += generate_settings({"MP2MIN": float(self.mp2min), "MP2MAX": float(self.mp2max), "CMIN": float(self.cmin),
                                        "CMAX": float(self.cmax)}, ["MP2MIN", "MP2MAX", "CMIN", "CMAX"])
        if self.laplace is not None:
            s += generate_settings({"BNODES": int(self.bnodes), "divergence_cleaning_form": "USE_LAPLACE_" + self.laplace},
                                   ["BNODES", "divergence_cleaning_form"])
        s += generate_settings({"periodic": "USE_PERIODIC" if self.periodic else "NONE"}, ["periodic"])
        if self.testcase == "uniform_flow":
            s += generate_settings({"UX": float(self.ux), "UY": float(self.uy), "UZ": float(self.uz)}, ["UX", "UY", "UZ"])
        return s

    @classmethod
    def from_settings(cls, settings: str, modules: Sequence[str]) -> "ProgramConfig":
        """Parse a settings string + module list back into a config."""
        defs: Dict[str, str] = {}
        for tok in settings.split():
            m = re.match(r"-D(\w+)(?:=(.*))?$", tok)
            if m:
                defs[m.group(1)] = m.group(2) if m.group(2) is not None else "1"
        base = [re.sub(r"\.(h|cl)$", "", m.split("/")[-1]) for m in modules]
        opset = next(m for m in base if m in OPSET_MODULES)
        testcase = next((m for m in base if m in TESTCASES), "uniform_flow")

        def pick(prefix, names, default):
            for n in names:
                if "%s%s" % (prefix, n) in defs:
                    return n
            return default

        diss = None
        if "USE_ARTIFICIAL_DISSIPATION" in defs:
            diss = pick("USE_", ["HIGH_ORDER_DISSIPATION", "FIRST_ORDER_DISSIPATION", "ADAPTIVE_DISSIPATION"], None)
            diss = diss.replace("_DISSIPATION", "") if diss else None
        lap = pick("USE_LAPLACE_", list(LAPLACE_FORMS), None) if "kernel_div_cleaning" in base else None
        f = lambda k, d: float(defs[k]) if k in defs else d  # noqa: E731
        i = lambda k, d: int(float(defs[k])) if k in defs else d  # noqa: E731
        return cls(real=defs.get("REAL", "double"), opset=opset, nx=i("NODES_X", 1), ny=i("NODES_Y", 1),
                   nz=i("NODES_Z", 1), xmin=f("XMIN", 0.0), xmax=f("XMAX", 1.0), ymin=f("YMIN", 0.0),
                   ymax=f("YMAX", 1.0), zmin=f("ZMIN", 0.0), zmax=f("ZMAX", 1.0), dx=f("DX", None), dy=f("DY", None),
                   dz=f("DZ", None), dt=f("DT", 0.0), periodic="USE_PERIODIC" in defs, hall="USE_HALL" in defs,
                   dissipation=diss, ho_factor=f("HO_DISSIPATION_FACTOR", 1.0), mp2min=f("MP2MIN", 1.0),
                   mp2max=f("MP2MAX", 10.0), cmin=f("CMIN", 1.0), cmax=f("CMAX", 10.0),
                   form_uibj=pick("USE_UIBJ_", UIBJ_FORMS, "PRODUCT"),
                   form_source=pick("USE_SOURCE_", SOURCE_FORMS, "CENTRAL"),
                   form_ujbi=pick("USE_UJBI_", UJBI_FORMS, "CENTRAL"), testcase=testcase, laplace=lap,
                   w_size=i("W_SIZE", 256), ux=f("UX", 0.0), uy=f("UY", 0.0), uz=f("UZ", 0.0))
