"""ORACLE (test infrastructure only): restatement of Kadath's 1-D spectral transforms.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use this
package, and only as the checker.  It is never the measured or shipped path.

Each function restates the reference routine cited in its docstring, step by
step, on a numpy array.  FFTW's r2hc / hc2r (the reference calls
fftw_plan_r2r_1d(n, ..., FFTW_R2HC / FFTW_HC2R), include/base_fftw.hpp:28-38) are
restated from FFTW's documented definitions with numpy.fft:
  r2hc:  out[k] = Re X_k (0 <= k <= n/2), out[n-k] = Im X_k (0 < k < n/2),
         X_k = sum_j x_j exp(-2 pi i j k / n)
  hc2r:  out[j] = sum_k Y_k exp(+2 pi i j k / n)  (unnormalised), Y Hermitian.

Pinned by: closed-form basis sampling (T_k on x_i = -cos(pi i/(N-1)), cos(2j theta),
sin((2j+1) theta), cos(m phi)/sin(m phi)), coef o coef_i = identity, and scipy's
DCT-I as an independent textbook routine (tests/test_oracle_spectral.py).
"""
import numpy as np

CHEB, CHEB_EVEN, CHEB_ODD, COSSIN, COS, COS_EVEN, COS_ODD = 1, 2, 3, 4, 5, 6, 7
SIN, SIN_EVEN, SIN_ODD, LEG, LEG_EVEN, LEG_ODD, COSSIN_EVEN, COSSIN_ODD = 8, 9, 10, 11, 12, 13, 14, 15


def r2hc(x):
    n = len(x)
    X = np.fft.rfft(x)
    out = np.empty(n)
    out[: n // 2 + 1] = X.real
    for k in range(1, (n + 1) // 2):
        out[n - k] = X[k].imag
    return out


def hc2r(hc):
    n = len(hc)
    Y = np.zeros(n // 2 + 1, dtype=complex)
    Y.real[:] = hc[: n // 2 + 1]
    for k in range(1, (n + 1) // 2):
        Y[k] += 1j * hc[n - k]
    return np.fft.irfft(Y, n) * n


def _fold_even(tab):
    """Folding shared by the cosine-type forward transforms (coef_1d.cpp:62-72)."""
    nr = len(tab)
    aux = np.zeros(nr - 1)
    for i in range(1, (nr - 1) // 2):
        fp = 0.5 * (tab[i] + tab[nr - 1 - i])
        fms = 0.5 * (tab[i] - tab[nr - 1 - i]) * np.sin(np.pi * i / (nr - 1))
        aux[i] = fp + fms
        aux[nr - 1 - i] = fp - fms
    aux[0] = 0.5 * (tab[0] + tab[nr - 1])
    aux[(nr - 1) // 2] = tab[(nr - 1) // 2]
    return aux


def _unfold_cos(tab_in, aux, fmoins0, last_div=True):
    """Even/odd coefficient recovery of coef_1d_cos_even (coef_1d.cpp:337-362)."""
    nr = len(tab_in)
    out = np.zeros(nr)
    out[0] = aux[0] / (nr - 1)
    for i in range(2, nr - 1, 2):
        out[i] = 2 * aux[i // 2] / (nr - 1)
    out[nr - 1] = aux[(nr - 1) // 2] / (nr - 1) if last_div else aux[(nr - 1) // 2]
    out[1] = 0
    som = 0.0
    for i in range(3, nr, 2):
        out[i] = out[i - 2] + 4 * aux[nr - 1 - i // 2] / (nr - 1)
        som += out[i]
    c1 = (fmoins0 - som) / ((nr - 1) // 2)
    out[1] = c1
    for i in range(3, nr, 2):
        out[i] += c1
    return out


def coef_1d_cheb(tab):
    """src/Coef/coef_1d.cpp:57-100."""
    tab = np.asarray(tab, dtype=float)
    nr = len(tab)
    fmoins0 = 0.5 * (tab[0] - tab[nr - 1])
    aux = _fold_even(tab)
    aux = r2hc(aux)
    out = np.zeros(nr)
    out[0] = aux[0] / (nr - 1)
    for i in range(2, nr - 1, 2):
        out[i] = 2 * aux[i // 2] / (nr - 1)
    out[nr - 1] = aux[(nr - 1) // 2] / (nr - 1)
    out[1] = 0
    som = 0.0
    for i in range(3, nr, 2):
        out[i] = out[i - 2] - 4 * aux[nr - 1 - (i - 1) // 2] / (nr - 1)
        som += out[i]
    c1 = -(fmoins0 + som) / ((nr - 1) // 2)
    out[1] = c1
    for i in range(3, nr, 2):
        out[i] += c1
    return out


def coef_1d_cos_even(tab):
    """src/Coef/coef_1d.cpp:314-363 (nbr > 3)."""
    tab = np.asarray(tab, dtype=float)
    nr = len(tab)
    if nr <= 3:
        return tab.copy()
    fmoins0 = 0.5 * (tab[0] - tab[nr - 1])
    aux = r2hc(_fold_even(tab))
    return _unfold_cos(tab, aux, fmoins0)


def coef_1d_cos_odd(tab):
    """src/Coef/coef_1d.cpp:365-414: multiply by cos(theta), cos_even transform,
    then cos(2j t) -> cos((2j+1) t) by the recurrence cf_i = 2 cf_i - cf_{i-1}."""
    tab = np.asarray(tab, dtype=float)
    nr = len(tab)
    if nr <= 3:
        return tab.copy()
    cf = np.zeros(nr)
    for i in range(nr - 1):
        cf[i] = tab[i] * np.sin(np.pi * (nr - 1 - i) / 2.0 / (nr - 1))
    cf[nr - 1] = 0
    fmoins0 = 0.5 * (cf[0] - cf[nr - 1])
    aux = r2hc(_fold_even(cf))
    cf = _unfold_cos(cf, aux, fmoins0, last_div=False)
    cf[0] = 2 * cf[0]
    for i in range(1, nr - 1):
        cf[i] = 2 * cf[i] - cf[i - 1]
    cf[nr - 1] = 0
    return cf


def coef_1d_sin_even(tab):
    """src/Coef/coef_1d.cpp:416-448 (nbr > 3)."""
    tab = np.asarray(tab, dtype=float)
    nr = len(tab)
    if nr <= 3:
        return tab.copy()
    aux = np.zeros(nr - 1)
    for i in range(1, (nr - 1) // 2):
        fp = 0.5 * (tab[i] + tab[nr - 1 - i]) * np.sin(np.pi * i / (nr - 1))
        fms = 0.5 * (tab[i] - tab[nr - 1 - i])
        aux[i] = fp + fms
        aux[nr - 1 - i] = fp - fms
    aux[0] = 0.5 * (tab[0] - tab[nr - 1])
    aux[(nr - 1) // 2] = tab[(nr - 1) // 2]
    aux = r2hc(aux)
    out = np.zeros(nr)
    out[0] = 0.0
    for i in range(2, nr - 1, 2):
        out[i] = -2 * aux[nr - 1 - i // 2] / (nr - 1)
    out[nr - 1] = 0
    out[1] = 2 * aux[0] / (nr - 1)
    for i in range(3, nr, 2):
        out[i] = out[i - 2] + 4 * aux[i // 2] / (nr - 1)
    return out


def coef_1d_sin_odd(tab):
    """src/Coef/coef_1d.cpp:450-499: multiply by sin(theta), cos_even transform,
    then cos(2j t) -> sin((2j+1) t) by cf_i = 2 cf_i + cf_{i-1}."""
    tab = np.asarray(tab, dtype=float)
    nr = len(tab)
    if nr <= 3:
        return tab.copy()
    cf = np.zeros(nr)
    for i in range(1, nr):
        cf[i] = tab[i] * np.sin(np.pi * i / 2.0 / (nr - 1))
    fmoins0 = 0.5 * (cf[0] - cf[nr - 1])
    aux = r2hc(_fold_even(cf))
    cf = _unfold_cos(cf, aux, fmoins0)
    cf[0] = 2 * cf[0]
    for i in range(1, nr - 1):
        cf[i] = 2 * cf[i] + cf[i - 1]
    cf[nr - 1] = 0
    return cf


def coef_1d_cossin(tab_conf):
    """src/Coef/coef_1d.cpp:254-292.  Input: n_phi values; output n_phi + 2
    coefficients [a0, 0, a1, b1, ..., a_{n/2}, 0]."""
    x = np.asarray(tab_conf, dtype=float)
    npt = len(x)
    nbr = npt + 2
    cf = r2hc(x)
    out = np.zeros(nbr)
    out[0] = cf[0] / npt
    out[1] = 0
    idx = 2
    pc, ps = 1, npt - 1
    for _ in range(1, nbr // 2 - 1):
        out[idx] = 2.0 * cf[pc] / npt
        idx += 1
        pc += 1
        out[idx] = -2.0 * cf[ps] / npt
        idx += 1
        ps -= 1
    out[idx] = cf[pc] / npt
    out[idx + 1] = 0
    return out


def _cos_even_synth(tab):
    """hc2r-based synthesis shared by coef_i_1d_cheb / cos_even (coef_i_1d.cpp:49-90, 310-349)."""
    nbr = len(tab)
    cf = np.zeros(nbr)
    aux = np.zeros(nbr - 1)
    c1 = tab[1]
    somme = 0.0
    for i in range(3, nbr, 2):
        cf[i] = tab[i] - c1
        somme += cf[i]
    for i in range(3, nbr, 2):
        aux[nbr - 1 - i // 2] = 0.25 * (cf[i] - cf[i - 2])
    aux[0] = tab[0]
    for i in range(1, (nbr - 1) // 2):
        aux[i] = 0.5 * tab[2 * i]
    aux[(nbr - 1) // 2] = tab[nbr - 1]
    aux = hc2r(aux)
    return aux, c1, somme


def coef_i_1d_cheb(tab):
    """src/Coef/coef_i_1d.cpp:49-90 (nbr > 3)."""
    tab = np.asarray(tab, dtype=float)
    nbr = len(tab)
    if nbr <= 3:
        return tab.copy()
    cf = np.zeros(nbr)
    aux = np.zeros(nbr - 1)
    c1 = tab[1]
    somme = 0.0
    for i in range(3, nbr, 2):
        cf[i] = tab[i] - c1
        somme += cf[i]
    fmoins0 = -((nbr - 1) // 2) * c1 - somme
    for i in range(3, nbr, 2):
        aux[nbr - 1 - i // 2] = -0.25 * (cf[i] - cf[i - 2])
    aux[0] = tab[0]
    for i in range(1, (nbr - 1) // 2):
        aux[i] = 0.5 * tab[2 * i]
    aux[(nbr - 1) // 2] = tab[nbr - 1]
    aux = hc2r(aux)
    out = np.zeros(nbr)
    for i in range(1, (nbr - 1) // 2):
        fp = 0.5 * (aux[i] + aux[nbr - 1 - i])
        fm = 0.5 * (aux[i] - aux[nbr - 1 - i]) / np.sin(np.pi * i / (nbr - 1))
        out[i] = fp + fm
        out[nbr - i - 1] = fp - fm
    out[0] = aux[0] + fmoins0
    out[nbr - 1] = aux[0] - fmoins0
    out[(nbr - 1) // 2] = aux[(nbr - 1) // 2]
    return out


def coef_i_1d_cos_even(tab):
    """src/Coef/coef_i_1d.cpp:310-349."""
    tab = np.asarray(tab, dtype=float)
    nbr = len(tab)
    aux, c1, somme = _cos_even_synth(tab)
    fmoins0 = ((nbr - 1) // 2) * c1 + somme
    out = np.zeros(nbr)
    for i in range(1, (nbr - 1) // 2):
        fp = 0.5 * (aux[i] + aux[nbr - 1 - i])
        fm = 0.5 * (aux[i] - aux[nbr - 1 - i]) / np.sin(np.pi * i / (nbr - 1))
        out[i] = fp + fm
        out[nbr - i - 1] = fp - fm
    out[0] = aux[0] + fmoins0
    out[nbr - 1] = aux[0] - fmoins0
    out[(nbr - 1) // 2] = aux[(nbr - 1) // 2]
    return out


def coef_i_1d_cos_odd(tab):
    """src/Coef/coef_i_1d.cpp:351-396."""
    tab = np.asarray(tab, dtype=float)
    nbr = len(tab)
    ti = np.zeros(nbr)
    ti[0] = 0.5 * tab[0]
    for i in range(1, nbr - 1):
        ti[i] = 0.5 * (tab[i] + tab[i - 1])
    ti[nbr - 1] = 0.5 * tab[nbr - 2]
    aux, c1, somme = _cos_even_synth(ti)
    fmoins0 = ((nbr - 1) // 2) * c1 + somme
    out = np.zeros(nbr)
    for i in range(1, (nbr - 1) // 2):
        fp = 0.5 * (aux[i] + aux[nbr - 1 - i])
        fm = 0.5 * (aux[i] - aux[nbr - 1 - i]) / np.sin(np.pi * i / (nbr - 1))
        out[i] = (fp + fm) / np.sin(np.pi * (nbr - 1 - i) / 2 / (nbr - 1))
        out[nbr - i - 1] = (fp - fm) / np.sin(np.pi * i / 2 / (nbr - 1))
    out[0] = aux[0] + fmoins0
    out[nbr - 1] = 0
    out[(nbr - 1) // 2] = aux[(nbr - 1) // 2] / np.sin(np.pi * (nbr - 1) / 4 / (nbr - 1))
    return out


def coef_i_1d_sin_even(tab):
    """src/Coef/coef_i_1d.cpp:398-427."""
    tab = np.asarray(tab, dtype=float)
    nbr = len(tab)
    aux = np.zeros(nbr - 1)
    for i in range(2, nbr - 1, 2):
        aux[nbr - 1 - i // 2] = -0.5 * tab[i]
    aux[0] = 0.5 * tab[1]
    for i in range(1, (nbr - 1) // 2):
        aux[i] = 0.25 * (tab[2 * i + 1] - tab[2 * i - 1])
    aux[(nbr - 1) // 2] = -0.5 * tab[nbr - 2]
    aux = hc2r(aux)
    out = np.zeros(nbr)
    for i in range(1, (nbr - 1) // 2):
        fp = 0.5 * (aux[i] + aux[nbr - 1 - i]) / np.sin(np.pi * i / (nbr - 1))
        fm = 0.5 * (aux[i] - aux[nbr - 1 - i])
        out[i] = fp + fm
        out[nbr - i - 1] = fp - fm
    out[0] = 0
    out[nbr - 1] = -2 * aux[0]
    out[(nbr - 1) // 2] = aux[(nbr - 1) // 2]
    return out


def coef_i_1d_sin_odd(tab):
    """src/Coef/coef_i_1d.cpp:429-475."""
    tab = np.asarray(tab, dtype=float)
    nbr = len(tab)
    ti = np.zeros(nbr)
    ti[0] = 0.5 * tab[0]
    for i in range(1, nbr - 1):
        ti[i] = 0.5 * (tab[i] - tab[i - 1])
    ti[nbr - 1] = -0.5 * tab[nbr - 2]
    aux, c1, somme = _cos_even_synth(ti)
    fmoins0 = ((nbr - 1) // 2) * c1 + somme
    out = np.zeros(nbr)
    for i in range(1, (nbr - 1) // 2):
        fp = 0.5 * (aux[i] + aux[nbr - 1 - i])
        fm = 0.5 * (aux[i] - aux[nbr - 1 - i]) / np.sin(np.pi * i / (nbr - 1))
        out[i] = (fp + fm) / np.sin(np.pi * i / 2 / (nbr - 1))
        out[nbr - i - 1] = (fp - fm) / np.sin(np.pi * (nbr - 1 - i) / 2 / (nbr - 1))
    out[0] = 0
    out[nbr - 1] = aux[0] - fmoins0
    out[(nbr - 1) // 2] = aux[(nbr - 1) // 2] / np.sin(np.pi * (nbr - 1) / 4 / (nbr - 1))
    return out


def coef_i_1d_cossin(tab):
    """src/Coef/coef_i_1d.cpp:217-239.  n_phi + 2 coefficients -> n_phi values."""
    tab = np.asarray(tab, dtype=float)
    nbr = len(tab)
    npt = nbr - 2
    cf = np.zeros(npt)
    cf[0] = tab[0]
    for i in range(1, npt // 2):
        cf[i] = 0.5 * tab[2 * i]
        cf[npt - i] = -0.5 * tab[2 * i + 1]
    cf[npt // 2] = tab[npt]
    return hc2r(cf)


COEF_1D = {CHEB: coef_1d_cheb, COS_EVEN: coef_1d_cos_even, COS_ODD: coef_1d_cos_odd, SIN_EVEN: coef_1d_sin_even,
           SIN_ODD: coef_1d_sin_odd, COSSIN: coef_1d_cossin}
COEF_I_1D = {CHEB: coef_i_1d_cheb, COS_EVEN: coef_i_1d_cos_even, COS_ODD: coef_i_1d_cos_odd,
             SIN_EVEN: coef_i_1d_sin_even, SIN_ODD: coef_i_1d_sin_odd, COSSIN: coef_i_1d_cossin}


def coef_1d(base, tab):
    """Dispatch table src/Coef/coef_1d.cpp:607-636."""
    return COEF_1D[base](tab)


def coef_i_1d(base, tab):
    """Dispatch table src/Coef/coef_i_1d.cpp:527-556."""
    return COEF_I_1D[base](tab)


def coef_3d(conf, base_r, base_t, base_p):
    """Base_spectral::coef (src/Coef/coef.cpp:96-104): phi, then theta (basis per
    phi coefficient k), then r (basis per (j, k)).  conf[k, j, i] (phi, theta, r)."""
    npt, nt, nr = conf.shape
    a = np.zeros((npt + 2, nt, nr))
    for j in range(nt):
        for i in range(nr):
            a[:, j, i] = coef_1d(base_p, conf[:, j, i])
    for k in range(npt + 2):
        for i in range(nr):
            a[k, :, i] = coef_1d(base_t[k], a[k, :, i])
    for k in range(npt + 2):
        for j in range(nt):
            a[k, j, :] = coef_1d(base_r[k * nt + j], a[k, j, :])
    return a


def coef_i_3d(coef, base_r, base_t, base_p):
    """Base_spectral::coef_i (src/Coef/coef_i.cpp:96-104): r, theta, phi."""
    npc, nt, nr = coef.shape
    a = coef.astype(float).copy()
    for k in range(npc):
        for j in range(nt):
            a[k, j, :] = coef_i_1d(base_r[k * nt + j], a[k, j, :])
    for k in range(npc):
        for i in range(nr):
            a[k, :, i] = coef_i_1d(base_t[k], a[k, :, i])
    out = np.zeros((npc - 2, nt, nr))
    for j in range(nt):
        for i in range(nr):
            out[:, j, i] = coef_i_1d(base_p, a[:, j, i])
    return out


def std_base(nt, np_phi, anti=False):
    """Domain_shell::set_cheb_base / set_anti_cheb_base (src/Domain/Spheric/domain_shell.cpp:288-309, 319-340):
    theta basis per phi coefficient k (m = k//2): COS_EVEN/SIN_ODD (or COS_ODD/SIN_EVEN
    for the z-antisymmetric base), r basis CHEB.  Returns (base_r[k*nt+j], base_t[k], base_p)."""
    npc = np_phi + 2
    bt = []
    for k in range(npc):
        m = k // 2
        if not anti:
            bt.append(COS_EVEN if m % 2 == 0 else SIN_ODD)
        else:
            bt.append(COS_ODD if m % 2 == 0 else SIN_EVEN)
    br = [CHEB] * (npc * nt)
    return np.array(br, dtype=np.int32), np.array(bt, dtype=np.int32), COSSIN
