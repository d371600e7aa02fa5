"""Kernel schedules of one time step, per integrator.

Restates the ``RK_Step`` kernel lists of
``matlab/+induction_eq/compute_numerical_solution_setup.m:140-371``: for every stage the
``calc_curlB_rho_*`` kernel of the field the stage differentiates (Hall only), then
``calc_u_*`` (time-dependent velocity only), then the stage kernels, and the closing
``calc_time_*`` kernel.  ``num_fields`` is 3 for the SSPRK and 2R methods (b1, b2, b3/db)
and 2 for the 2N methods (b1, b2).
"""
from __future__ import annotations

from typing import List, Tuple

INTEGRATORS = ["SSPRK33", "SSPRK104", "KennedyCarpenterLewis2R54C", "CalvoFrancoRandez2R64",
               "CarpenterKennedy2N54", "ToulorgeDesmet2N84F"]

# number of right-hand-side evaluations per step
STAGES_PER_STEP = {"SSPRK33": 3, "SSPRK104": 10, "KennedyCarpenterLewis2R54C": 5, "CalvoFrancoRandez2R64": 6,
                   "CarpenterKennedy2N54": 5, "ToulorgeDesmet2N84F": 8}


def rk_schedule(integrator: str, hall: bool, variable_u: bool) -> Tuple[List[str], int]:
    """Return (kernel names of one step, number of field registers)."""
    s: List[str] = []

    def pre(curl_name: str, u_name: str):
        if hall:
            s.append(curl_name)
        if variable_u:
            s.append(u_name)

    if integrator == "SSPRK33":
        pre("calc_curlB_rho_1_3args", "calc_u_3_args")
        s.append("SSPRK33_1")
        s.append("SSPRK33_2a")
        pre("calc_curlB_rho_2_3args", "calc_u_3_args")
        s.append("SSPRK33_2b")
        s.append("SSPRK33_3a")
        pre("calc_curlB_rho_3_3args", "calc_u_3_args")
        s.append("SSPRK33_3b")
        s.append("calc_time_3_args")
        return s, 3
    if integrator == "SSPRK104":
        # (time kernel, curl field index, stage kernel); None time kernel = none
        plan = [(None, 1, "SSPRK104_01"), ("SSPRK104_02a", 2, "SSPRK104_02b"), ("SSPRK104_03a", 3, "SSPRK104_03b"),
                ("SSPRK104_04a", 2, "SSPRK104_04b"), ("SSPRK104_05a", 3, "SSPRK104_05b"),
                ("SSPRK104_06", 2, "SSPRK104_07"), ("SSPRK104_08a", 1, "SSPRK104_08b"),
                ("SSPRK104_09a", 2, "SSPRK104_09b"), ("SSPRK104_10a", 1, "SSPRK104_10b"),
                ("SSPRK104_11a", 2, "SSPRK104_11b")]
        for tk, f, sk in plan:
            if tk:
                s.append(tk)
            pre("calc_curlB_rho_%d_3args" % f, "calc_u_3_args")
            s.append(sk)
        s.append("calc_time_3_args")
        return s, 3
    if integrator == "KennedyCarpenterLewis2R54C":
        a = ["KennedyCarpenterLewis2R54C_1a", "KennedyCarpenterLewis2R54C_2a", "KennedyCarpenterLewis2R54C_3a",
             "KennedyCarpenterLewis2R54C_4a", "KennedyCarpenterLewis2R54C_5"]
        b = ["KennedyCarpenterLewis2R54C_1b", "KennedyCarpenterLewis2R54C_2b", "KennedyCarpenterLewis2R54C_3b",
             "KennedyCarpenterLewis2R54C_4b", "calc_time_3_args"]
        curl = [1, 1, 2, 1, 2]
        for ka, kb, f in zip(a, b, curl):
            pre("calc_curlB_rho_%d_3args" % f, "calc_u_3_args")
            s += [ka, kb]
        return s, 3
    if integrator == "CalvoFrancoRandez2R64":
        a = ["CalvoFrancoRandez2R64_%da" % i for i in range(1, 6)] + ["CalvoFrancoRandez2R64_6"]
        b = ["CalvoFrancoRandez2R64_%db" % i for i in range(1, 6)] + ["calc_time_3_args"]
        curl = [1, 2, 2, 2, 2, 2]
        for ka, kb, f in zip(a, b, curl):
            pre("calc_curlB_rho_%d_3args" % f, "calc_u_3_args")
            s += [ka, kb]
        return s, 3
    if integrator in ("CarpenterKennedy2N54", "ToulorgeDesmet2N84F"):
        n = 5 if integrator == "CarpenterKennedy2N54" else 8
        for i in range(1, n + 1):
            pre("calc_curlB_rho_1_2args", "calc_u_2_args")
            s += ["%s_%da" % (integrator, i), "%s_%db" % (integrator, i)]
        s.append("calc_time_2_args")
        return s, 2
    raise ValueError("Unkown time integrator %s!" % integrator)
