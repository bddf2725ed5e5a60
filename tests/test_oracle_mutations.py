"""The oracle's pins catch plausible restatement mistakes.

Each case plants one mistake into a copy of oracle/pfem_oracle.cpp (on the CPU, in a temporary
directory), builds it, and runs the pin tests against that build: at least one pin must fail.
The mistakes are the slips a restatement of the reference is most likely to make — a
flipped sign, swapped factors, a dropped square, the wrong stabilisation or deviatoric
matrix, an index or denominator mix-up.
"""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "pfem_oracle.cpp")
PINS = ["tests/test_oracle_pins.py", "tests/test_oracle_wc_pins.py"]

MUTATIONS = {
    # WC momentum: sign of D^T P (MomEquation.inl:229-302)
    "wc_dtp_sign": ("Fte[e * nd + r] = -kv + dp + Fe[r];", "Fte[e * nd + r] = -kv - dp + Fe[r];"),
    # WC continuity: K0 and K0' swapped in the D factor (ContEquation.inl:80-86)
    "wc_k0_swap": ("return W.K0 + W.K0p * pg;", "return W.K0p + W.K0 * pg;"),
    # WC continuity: Meduri and lumped stabilisation swapped (ContEquation.inl:353-413)
    "wc_meduri_swap": ("if (W.meduri)", "if (!W.meduri)"),
    # WC deviatoric ddev replaced by the incompressible one (MomEquation.inl:63-83)
    "wc_ddev": ("dd[r][c] = (r == c) ? 4.0 / 3 : -2.0 / 3;", "dd[r][c] = (r == c) ? 2.0 : 0.0;"),
    # WC body force without the density factor (MomEquation.inl:135-139)
    "wc_force_rho": ("mb.Fv(geo.detJ, rhoG, W.g, Fe);", "mb.Fv(geo.detJ, [](const double*) { return 1000.0; }, W.g, Fe);"),
    # CFL: viscous term without its square (WCompNewton/Solver.cpp:214-217)
    "cfl_alpha_square": ("maxS = std::max(maxS, 4 * alpha * alpha / (he * he));",
                         "maxS = std::max(maxS, 4 * alpha / (he * he));"),
    # CFL: c^2 without the K0' p term (ContEquation.inl:117-121)
    "cfl_c2": ("const double c2 = (W.K0 + W.K0p * p[nn]) / rho[nn];", "const double c2 = W.K0 / rho[nn];"),
    # Tait EOS exponent (ContEquation.inl:319-332)
    "tait_exponent": ("rho[nn] = std::pow((W.K0p / W.K0) * p[nn] + 1, 1 / W.K0p) * W.rhoStar;",
                      "rho[nn] = std::pow((W.K0p / W.K0) * p[nn] + 1, W.K0p) * W.rhoStar;"),
    # Picard U residual: iterate instead of previous iterate in the denominator (PSPG.inl:333)
    "res_den": ("const double a = qIter[nn + d * N], b = qIterPrev[nn + d * N];\n                num += (a - b) * (a - b);\n                den += b * b;",
                "const double a = qIter[nn + d * N], b = qIterPrev[nn + d * N];\n                num += (a - b) * (a - b);\n                den += a * a;"),
    # PSPG tau: viscous term without the factor 9 (PSPG.inl:257-258)
    "tau_nine": ("9 * (4 * P.mu / (h * h * P.rho))", "(4 * P.mu / (h * h * P.rho))"),
    # PSPG free-node rhs without dt g (PSPG.inl:193-204)
    "bc_free_rhs": ("S.b[nn + d * N] = qPrev[nn + d * N] + P.dt * P.g[d];", "S.b[nn + d * N] = qPrev[nn + d * N];"),
}

HEAT_PINS = ["tests/test_oracle_heat_pins.py"]
FRAC_PINS = ["tests/test_oracle_frac_pins.py"]
BQ_PINS = ["tests/test_oracle_boussinesq_pins.py"]
FST_PINS = ["tests/test_oracle_fst_pins.py"]
# (old, new, pin files) of the heat and fractional-step restatements
MUTATIONS_PINNED = {
    # heat: L without the time step (HeatEquation.inl:245-250)
    "heat_dt_L": ("for (int j = 0; j < ne; ++j) Le[i][j] = P.dt * Le[i][j];",
                  "for (int j = 0; j < ne; ++j) Le[i][j] = Le[i][j];", HEAT_PINS),
    # heat: convection added instead of subtracted (HeatEquation.inl:331-341)
    "heat_conv_sign": ("for (int i = 0; i < D; ++i) S.b[fn[i]] -= P.dt * sgh[i];",
                       "for (int i = 0; i < D; ++i) S.b[fn[i]] += P.dt * sgh[i];", HEAT_PINS),
    # heat: radiation with T^3 (HeatEquation.inl:95-99)
    "heat_rad_power": ("(Tg * Tg * Tg * Tg - P.Tinf * P.Tinf * P.Tinf * P.Tinf)",
                       "(Tg * Tg * Tg - P.Tinf * P.Tinf * P.Tinf)", HEAT_PINS),
    # FracStep: M instead of M/dt in the prediction matrix and right-hand side (FracStep.inl:43)
    "frac_m_dt": ("Ms_dt[a][b] = idt * Ms[a][b];", "Ms_dt[a][b] = Ms[a][b];", FRAC_PINS),
    # FracStep: sign of gammaFS D^T p_prev (FracStep.inl:61)
    "frac_dtp_sign": ("DTp[r] = gammaFS * t;", "DTp[r] = -gammaFS * t;", FRAC_PINS),
    # FracStep: rho dt instead of rho / dt in the pressure right-hand side (FracStep.inl:292)
    "frac_pcorr_scale": ("-((P.rho / P.dt) * Dv)", "-((P.rho * P.dt) * Dv)", FRAC_PINS),
    # FracStep: gammaFS dropped from L p_prev (FracStep.inl:293)
    "frac_pcorr_gamma": ("{nodes[i], gammaFS * Lp}", "{nodes[i], Lp}", FRAC_PINS),
    # FracStep: dp = p - p_prev whatever gammaFS (FracStep.inl:533)
    "frac_dp_gamma": ("dP[nn] = pNew[nn] - gammaFS * pPrev[nn];", "dP[nn] = pNew[nn] - pPrev[nn];", FRAC_PINS),
    # FracStep: sign of the velocity correction (FracStep.inl:375-380)
    "frac_vstep_sign": ("{nodes[i] + d * N, P.dt * t}", "{nodes[i] + d * N, -P.dt * t}", FRAC_PINS),
    # FracStep: K coupling block (d, d2) read as (d, d) (FracStep.inl:108-112)
    "frac_k_block": ("Ke[i + d * ne][j + d2 * ne]};", "Ke[i + d * ne][j + d * ne]};", FRAC_PINS),
    # Boussinesq: F factor with the sign of alpha flipped (MomContEquation.inl:170)
    "bq_f_sign": ("bq ? P.rho * (1 - P.mat.alpha * (Tg(Ng) - P.mat.Tr)) : P.rho",
                  "bq ? P.rho * (1 + P.mat.alpha * (Tg(Ng) - P.mat.Tr)) : P.rho", BQ_PINS),
    # Boussinesq: H factor left at 1 (MomContEquation.inl:191-196)
    "bq_h_factor": ("return bq ? 1 - P.mat.alpha * (Tg(Ng) - P.mat.Tr) : 1.0;", "return 1.0;", BQ_PINS),
    # FST: a wrong entry of the 3D T(P) (MatricesBuilder.inl:205)
    "fst_t_entry": ("T[0][1] = P[3] * P[3];", "T[0][1] = P[3] * P[4];", FST_PINS),
    # FST: sign of the 2D shear projector entry (MatricesBuilder.inl:176)
    "fst_p_shear": ("Pv[2] = -n[0] * n[1];", "Pv[2] = n[0] * n[1];", FST_PINS),
    # FST: without the reference element size (MatricesBuilder.inl:400)
    "fst_refsize": ("for (int c = 0; c < nd; ++c) FST[c] *= detJ * q.refSize;",
                    "for (int c = 0; c < nd; ++c) FST[c] *= detJ;", FST_PINS),
    # FST: the WC every-node rule used for PSPG too (MomContEquationPSPG.inl:162-171)
    "fst_any_node": ("if (allNodes ? nfs < D : nfs == 0) continue;", "if (nfs < D) continue;", FST_PINS),
    # FracStep: free nodes keep v_prev without dt g (FracStep.inl:244-252)
    "frac_free_rhs": ("S.b[nn + d * N] = vPrevQ[nn + d * N] + dt * P.g[d];", "S.b[nn + d * N] = vPrevQ[nn + d * N];",
                      FRAC_PINS),
}


@pytest.mark.parametrize("name", sorted(MUTATIONS) + sorted(MUTATIONS_PINNED))
def test_planted_mistake_fails_a_pin(name, tmp_path):
    old, new, pins = MUTATIONS[name] + (PINS,) if name in MUTATIONS else MUTATIONS_PINNED[name]
    src = open(SRC).read()
    assert src.count(old) == 1, f"mutation site not found once: {old}"
    mut = tmp_path / "pfem_oracle.cpp"
    mut.write_text(src.replace(old, new))
    lib = tmp_path / "liboracle_mut.so"
    subprocess.run(["g++", "-O1", "-std=c++17", "-fopenmp", "-fPIC", "-shared", "-o", str(lib), str(mut)], check=True,
                   capture_output=True)
    env = dict(os.environ, PFEM_ORACLE_LIB=str(lib), OMP_NUM_THREADS="2")
    r = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider", *pins], cwd=ROOT, env=env,
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 1, f"no pin caught the mutation {name}:\n{r.stdout[-2000:]}"
