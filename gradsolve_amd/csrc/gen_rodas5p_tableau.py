"""Writes gsx_rodas5p_tab.h, the Rodas5P tableau in the alpha/Gamma form the kernels and the
oracle use (``python gradsolve_amd/csrc/gen_rodas5p_tableau.py``; the header is committed and a
CPU test checks it against this script's output).

Rodas5P is G. Steinebach's order-5(4) Rosenbrock-Wanner pair with an order-4 continuous
extension (BIT Numer. Math. 63:27, 2023, doi:10.1007/s10543-023-00967-x). The coefficients
below are the published ones in the stored a/C form (gradsolve/solvers/rodas5p_step.py:32-97
lists the same table). The stored form converts to the step's alpha/Gamma form by
    Gamma^-1 = I / gamma - C,   alpha = a Gamma,   m = b Gamma,   m - mhat = btilde Gamma,
with b = (last row of a, 1) and btilde = e_8 (the pair is stiffly accurate), and the dense
weights for the k_i increments are H Gamma. The conversion runs in NumPy float64 with LAPACK's
inverse, as the reference does at import (rodas5p_step.py:63-75), so the constants round the
same way.
"""
import os

import numpy as np

S = 8
GAMMA_DIAG = 0.21193756319429014
A = np.array([
    [0, 0, 0, 0, 0, 0, 0, 0],
    [3.0, 0, 0, 0, 0, 0, 0, 0],
    [2.849394379747939, 0.45842242204463923, 0, 0, 0, 0, 0, 0],
    [-6.954028509809101, 2.489845061869568, -10.358996098473584, 0, 0, 0, 0, 0],
    [2.8029986275628964, 0.5072464736228206, -0.3988312541770524, -0.04721187230404641, 0, 0, 0,
     0],
    [-7.502846399306121, 2.561846144803919, -11.627539656261098, -0.18268767659942256,
     0.030198172008377946, 0, 0, 0],
    [-7.502846399306121, 2.561846144803919, -11.627539656261098, -0.18268767659942256,
     0.030198172008377946, 1, 0, 0],
    [-7.502846399306121, 2.561846144803919, -11.627539656261098, -0.18268767659942256,
     0.030198172008377946, 1, 1, 0],
])
C7 = np.array([
    [0, 0, 0, 0, 0, 0, 0],
    [-14.155112264123755, 0, 0, 0, 0, 0, 0],
    [-17.97296035885952, -2.859693295451294, 0, 0, 0, 0, 0],
    [147.12150275711716, -1.41221402718213, 71.68940251302358, 0, 0, 0, 0],
    [165.43517024871676, -0.4592823456491126, 42.90938336958603, -5.961986721573306, 0, 0, 0],
    [24.854864614690072, -3.0009227002832186, 47.4931110020768, 5.5814197821558125,
     -0.6610691825249471, 0, 0],
    [30.91273214028599, -3.1208243349937974, 77.79954646070892, 34.28646028294783,
     -19.097331116725623, -28.087943162872662, 0],
    [37.80277123390563, -3.2571969029072276, 112.26918849496327, 66.9347231244047,
     -40.06618937091002, -54.66780262877968, -9.48861652309627],
])
ABSCISSAE = np.array([0.0, 0.6358126895828704, 0.4095798393397535, 0.9769306725060716,
                      0.4288403609558664, 1.0, 1.0, 1.0])
GAMMA_SUMS = np.array([0.21193756319429014, -0.42387512638858027, -0.3384627126235924,
                       1.8046452872882734, 2.325825639765069, 0.0, 0.0, 0.0])
H = np.array([
    [25.948786856663858, -2.5579724845846235, 10.433815404888879, -2.3679251022685204,
     0.524948541321073, 1.1241088310450404, 0.4272876194431874, -0.17202221070155493],
    [-9.91568850695171, -0.9689944594115154, 3.0438037242978453, -24.495224566215796,
     20.176138334709044, 15.98066361424651, -6.789040303419874, -6.710236069923372],
    [11.419903575922262, 2.8879645146136994, 72.92137995996029, 80.12511834622643,
     -52.072871366152654, -59.78993625266729, -0.15582684282751913, 4.883087185713722],
])


def convert():
    """(alpha, Gamma, m, m - mhat, hd) of the alpha/Gamma form."""
    Cs = np.zeros((S, S))
    Cs[:, :7] = C7
    b = np.concatenate([A[7, :7], [1.0]])
    btilde = np.eye(S)[7]
    gam = np.linalg.inv(np.eye(S) / GAMMA_DIAG - Cs)
    alpha = A @ gam
    m = b @ gam
    mhat = m - btilde @ gam
    err = np.array([float(m[i] - mhat[i]) for i in range(S)])
    hd = H @ gam
    return alpha, gam, m, err, hd


def _arr(name, v):
    body = ", ".join(float(x).hex() for x in np.ravel(v))
    return f"constexpr double {name}[{np.size(v)}] = {{{body}}};\n"


def header() -> str:
    alpha, gam, m, err, hd = convert()
    out = ["// gsx_rodas5p_tab.h -- the Rodas5P tableau in alpha/Gamma form, written by\n",
           "// gen_rodas5p_tableau.py (Steinebach 2023, converted from the published a/C form\n",
           "// with NumPy float64). Do not edit by hand.\n",
           "#pragma once\n\n",
           "namespace gsx_rodas5p_tab {\n",
           f"constexpr int kStages = {S};\n",
           f"constexpr double kGamma = {GAMMA_DIAG.hex()};\n",
           "// alpha[i][j] (j < i): stage inputs y + sum_j alpha_ij k_j\n",
           _arr("kAlpha", alpha),
           "// Gamma[i][j] (j < i): the J-coupling sum_j Gamma_ij k_j of stage i\n",
           _arr("kGammaOff", np.tril(gam, -1)),
           "// c_i (time abscissae) and d_i = sum_j Gamma_ij (df/dt weights), as published\n",
           _arr("kC", ABSCISSAE),
           _arr("kD", GAMMA_SUMS),
           "// sum_j Gamma_ij of the converted Gamma (NumPy's row sum, diagonal included): the\n",
           "// df/dt weight of stage i for fields that depend on t\n",
           _arr("kGammaSum", gam.sum(axis=1)),
           "// m_i (solution weights) and m_i - mhat_i (error weights)\n",
           _arr("kM", m),
           _arr("kE", err),
           "// dense-output weights for the k_i: D_r = sum_i hd[r][i] k_i, r = 0, 1, 2\n",
           _arr("kHD", hd),
           "}  // namespace gsx_rodas5p_tab\n"]
    return "".join(out)


if __name__ == "__main__":
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "gsx_rodas5p_tab.h")
    with open(path, "w") as fh:
        fh.write(header())
    print(path)
