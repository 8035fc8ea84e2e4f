"""Oracle (test infrastructure only): CPU restatement of the VBRDF supplement
for the kernels the AeroApps MODIS land surface uses, the checker for the GPU
kernel k_brdf (aeroapps_amd/csrc/brdf.hip).  Nothing in the product imports it.

* getquad2            <- VB/vbrdf_sup_aux.f90:180-262 (Gauss-Legendre on [a, b])
* lambertian / rossthick / lisparse / coxmunk, derfc_e
                      <- VB/vbrdf_sup_kernels.f90:201-233, 287-332, 478-582,
                         1018-1141; VB/vbrdf_sup_aux.f90:266-286
* coxmunk_setup       <- V283/VLIDORT_SurfaceMod.F90:815-1042
                         (VLIDORT_CoxMunk_SingleGeom: wind speed -> sigma^2)
* vbrdf_master        <- VB/vbrdf_sup_masters.f90:1559-3293 (VBRDF_MAINMASTER)
                         with VB/vbrdf_sup_routines.f90:153-687 (VBRDF_MAKER),
                         :1962-2069 (SCALING_FOURIER_ZERO), :2073-2491
                         (VBRDF_FOURIER) and VB/vbrdf_sup_aux.f90:356-461
                         (BRDF_QUADRATURE_GAUSSIAN), restricted to what the
                         MODIS driver switches on: scalar, solar sources,
                         observation geometry (one geometry), half-range
                         azimuth quadrature, no emission, no glitter
                         corrections, white-sky albedo computed and checked.
* bpdf_ndvi, fresnel_vector
                      <- VB/vbrdf_sup_kernels.f90:3017-3198 (BPDFNDVI_VFUNCTION),
                         :3575-3702 (FRESNEL_VECTOR): the polarised land kernel,
                         3 x 3 Mueller block (checked against the compiled
                         reference's direct-bounce elements and closed forms)
* landmodis_setup     <- V283/VLIDORT_SurfaceMod.F90:1376-1596
                         (VLIDORT_LANDMODIS_SingleGeom: kernels Lambertian /
                         Ross-thick / Li-sparse, factors fiso / fvol / fgeo,
                         50 azimuth streams)

Sums run in the reference's order (DOT_PRODUCT left to right from zero,
kernels accumulated in order).  Pinned by the reference's own BRDF
self-tester results (vlidort_s_test/saved_results/gfortran/
results_brdf_self_tester.res, Ross-thick and Li-sparse blocks), by the
compiled reference (oracle/_ref, fixture tests/golden/brdf_modis.npz) and by
closed forms (tests/test_brdf.py).

Paths: VB = src/Components/rtms/RTSI/VLIDORT2p8p3/vsup/vbrdf,
V283 = src/Components/rtms/RTSI/VLIDORT2p8p3.
"""
from __future__ import annotations

import math

import numpy as np

# vlidort_pars.f90:391-407
DEG_TO_RAD = 1.7453292519943e-02
PIE = 180.0 * DEG_TO_RAD
PIO2 = 0.5 * PIE
PIO4 = 0.25 * PIE
MAXSTREAMS_SCALING = 24          # vlidort_pars.f90:232

LAMBERTIAN_IDX, ROSSTHICK_IDX, LISPARSE_IDX, COXMUNK_IDX = 1, 3, 4, 9


def getquad2(a: float, b: float, n: int):
    """Gauss-Legendre roots and weights on [a, b], Newton from the asymptotic guess."""
    roots = [0.0] * n
    wgts = [0.0] * n
    n2 = (n + 1) // 2
    nr = float(n)
    midpt = 0.5 * (b + a)
    sfac = 0.5 * (b - a)
    for m in range(1, n2 + 1):
        xx = PIE * (float(m) - 0.25) / (nr + 0.5)
        x = (1.0 - (nr - 1.0) / (8.0 * nr ** 3)
             - 1.0 / (384.0 * nr ** 4) * (39.0 - 28.0 / math.sin(xx) ** 2)) * math.cos(xx)
        while True:
            lpm1, lp = 0.0, 1.0
            for i in range(1, n + 1):
                ir = float(i)
                lpm2, lpm1 = lpm1, lp
                lp = ((2.0 * ir - 1.0) * x * lpm1 - (ir - 1.0) * lpm2) / ir
            dlp = nr * (x * lp - lpm1) / (x ** 2 - 1.0)
            xold = x
            x = xold - lp / dlp
            if abs(x - xold) <= 1.0e-13:
                break
        roots[m - 1] = midpt - sfac * x
        wgts[m - 1] = (2.0 * sfac) / ((1.0 - x ** 2) * dlp ** 2)
        roots[n - m] = midpt + sfac * x
        wgts[n - m] = wgts[m - 1]
    return roots, wgts


# Kernel functions: (xi, sxi) incident cosine/sine, (xj, sxj) reflected,
# phi relative azimuth [rad] with its cosine and sine.
def lambertian(pars, xi, sxi, xj, sxj, phi, cphi, sphi):
    return 1.0


def rossthick(pars, xi, sxi, xj, sxj, phi, cphi, sphi):
    ckphi = -cphi
    ds1 = xi * xj
    ds2 = sxi * sxj
    ds3 = xi + xj
    cksi = ds1 + ds2 * ckphi
    if cksi > 1.0:
        cksi = 1.0
    sksi = math.sqrt(1.0 - cksi * cksi)
    ksi = math.acos(cksi)
    func = ((PIO2 - ksi) * cksi + sksi) / ds3
    return func - PIO4


def lisparse(pars, xi, sxi, xj, sxj, phi, cphi, sphi):
    ckphi = -cphi
    t_inc = pars[1] * (sxi / xi)
    t_inc_sq = t_inc * t_inc
    ang = math.atan(t_inc)
    x_inc, sx_inc = math.cos(ang), math.sin(ang)
    t_ref = pars[1] * (sxj / xj)
    t_ref_sq = t_ref * t_ref
    ang = math.atan(t_ref)
    x_ref, sx_ref = math.cos(ang), math.sin(ang)
    cksi = x_inc * x_ref + sx_inc * sx_ref * ckphi
    p = (1.0 + cksi) / x_ref / x_inc
    a = 1.0 / x_inc
    b = 1.0 / x_ref
    r = a + b
    dt1 = t_ref_sq + t_inc_sq
    dt2 = t_inc * t_ref
    dt2sq = dt2 * dt2
    delta = math.sqrt(dt1 - 2.0 * dt2 * ckphi)
    dsq = delta * delta
    h = math.sqrt(dsq + sphi * sphi * dt2sq)
    cost = pars[0] * h / r
    if cost > 1.0:
        q = 1.0
    else:
        t = math.acos(cost)
        sint = math.sqrt(1.0 - cost * cost)
        q = 1.0 - ((t - sint * cost) / PIE)
    return 0.5 * p - q * r


def fresnel_vector(n1, n2, xi, sxi, xj, sxj, ckphi, skphi):
    """Fresnel amplitudes rotated to the meridian planes of the incident
    (azimuth 0) and reflected (azimuth phi) directions: (f11, f12, f21, f22,
    dmod), dmod the fourth power of |incident x reflected| (1 when they are
    parallel, with f11 = r_par, f22 = r_per)."""
    vi = (sxi * 1.0, sxi * 0.0, -xi)
    vr = (sxj * ckphi, sxj * skphi, xj)
    u = (vi[0] - vr[0], vi[1] - vr[1], vi[2] - vr[2])
    fac = math.sqrt(1.0 / (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]))
    ci = fac * (u[0] * vi[0] + u[1] * vi[1] + u[2] * vi[2])
    ct = math.sqrt(1.0 - (1.0 - ci * ci) * n1 * n1 / (n2 * n2))
    rper = (n1 * ci - n2 * ct) / (n1 * ci + n2 * ct)
    rpar = (n2 * ci - n1 * ct) / (n2 * ci + n1 * ct)
    ti = (-xi * 1.0, -xi * 0.0, -sxi)
    tr = (xj * ckphi, xj * skphi, -sxj)
    pi_ = (-0.0, 1.0, 0.0)
    pr = (-skphi, ckphi, 0.0)
    dot = lambda a, b: a[0] * b[0] + a[1] * b[1] + a[2] * b[2]
    pikr, prki, tikr, trki = dot(pi_, vr), dot(pr, vi), dot(ti, vr), dot(tr, vi)
    e1, e2, e3, e4 = pikr * prki, tikr * trki, tikr * prki, pikr * trki
    f11 = e1 * rper + e2 * rpar
    f12 = -(e3 * rper) + e4 * rpar
    f21 = -(e4 * rper) + e3 * rpar
    f22 = e2 * rper + e1 * rpar
    vp = (vi[1] * vr[2] - vi[2] * vr[1], vi[2] * vr[0] - vi[0] * vr[2], vi[0] * vr[1] - vi[1] * vr[0])
    dmod = dot(vp, vp) ** 2
    if dmod == 0.0:
        f11, f22, dmod = rpar, rper, 1.0
    return f11, f12, f21, f22, dmod


def bpdf_ndvi(pars, xi, sxi, xj, sxj, cphi, skphi):
    """3 x 3 Mueller block of the BPDF-NDVI kernel, pars = (refractive index,
    NDVI, scaling factor C)."""
    if abs(xi - 1.0) < 1e-9:
        xi = 0.999999999999
    if abs(xj - 1.0) < 1e-9:
        xj = 0.999999999999
    z = min(xi * xj - sxi * sxj * cphi, 1.0)
    cg = math.cos(math.acos(z) * 0.5)
    ndvi = pars[1] if -1.0 <= pars[1] <= 1.0 else 0.0
    fp0 = 0.25 * math.exp(-ndvi) / (xi + xj)
    att = math.exp(-math.sqrt(1.0 - cg * cg) / cg)
    f11, f12, f21, f22, dmod = fresnel_vector(1.0, pars[0], xi, sxi, xj, sxj, cphi, skphi)
    h = pars[2] * fp0 * att / dmod
    a11, a12, a21, a22 = f11 * f11, f12 * f12, f21 * f21, f22 * f22
    k = [[0.0] * 3 for _ in range(3)]
    k[0][0] = (a11 + a12 + a21 + a22) * 0.5 * h
    k[0][1] = (a11 - a12 + a21 - a22) * 0.5 * h
    k[1][0] = (a11 - a22 + a12 - a21) * 0.5 * h
    k[1][1] = (a11 - a12 - a21 + a22) * 0.5 * h
    k[0][2] = -(f11 * f12 + f21 * f22) * h
    k[1][2] = -(f11 * f12 - f21 * f22) * h
    k[2][0] = -(f11 * f21 + f12 * f22) * h
    k[2][1] = -(f11 * f21 - f12 * f22) * h
    k[2][2] = (f11 * f22 + f12 * f21) * h
    return k


def derfc_e(x: float) -> float:
    """The supplement's complementary error function (VB/vbrdf_sup_aux.f90:266-286),
    a Chebyshev fit with fractional error below 1.2e-7."""
    z = abs(x)
    t = 1.0 / (1.0 + 0.5 * z)
    r = t * math.exp(-z * z - 1.26551223 + t * (1.00002368 + t * (.37409196 + t * (.09678418 + t * (
        -.18628806 + t * (.27886807 + t * (-1.13520398 + t * (1.48851587 + t * (-.82215223 + t * .17087277)))))))))
    if x < 0.0:
        r = 2.0 - r
    return r


def coxmunk(pars, xi, sxi, xj, sxj, phi, ckphi, skphi):
    """COXMUNK_VFUNCTION (VB/vbrdf_sup_kernels.f90:1018-1141): pars = (sigma^2,
    m^2, shadow flag)."""
    ckphi_neg = -ckphi
    z = xi * xj + sxi * sxj * ckphi_neg
    if z > 1.0:
        z = 1.0
    z1 = math.acos(z)
    z2 = math.cos(z1 * 0.5)
    z2_sq_m1 = z2 * z2 + -1.0
    h1 = pars[1] * z2
    h2 = math.sqrt(pars[1] + z2_sq_m1)
    rp = (h1 - h2) / (h1 + h2)
    rl = (z2 - h2) / (z2 + h2)
    xmp = 0.5 * (rp * rp + rl * rl)
    a = 2.0 * z2
    b = (xi + xj) / a
    if b > 1.0:
        b = 1.0
    a = PIO2 - math.asin(b)
    ta = math.tan(a)
    argument = ta * ta / pars[0]
    k = 0.0
    if argument < 88.0:
        prob = math.exp(-argument)
        fac1 = prob / pars[0]
        fac2 = 0.25 / xi / (b ** 4)
        k = xmp * fac1 * fac2 / xj
    if pars[2] == 0.0:
        return k
    s1 = math.sqrt(pars[0] / PIE)
    s3 = 1.0 / math.sqrt(pars[0])
    s2 = s3 * s3
    shadow = []
    for x in (xi, xj):
        dcot = x / math.sqrt(1.0 - x * x) if x < 1.0 else math.inf
        t1 = math.exp(-dcot * dcot * s2)
        t2 = derfc_e(dcot * s3)
        shadow.append(0.5 * (s1 * t1 / dcot - t2))
    return k * (1.0 / (1.0 + shadow[0] + shadow[1]))


KERNELS = {LAMBERTIAN_IDX: lambertian, ROSSTHICK_IDX: rossthick, LISPARSE_IDX: lisparse, COXMUNK_IDX: coxmunk}


def _dot(a, b):
    s = 0.0
    for x, y in zip(a, b):
        s = s + x * y
    return s


def landmodis_setup(fiso: float, fgeo: float, fvol: float, param):
    """Kernel list of VLIDORT_LANDMODIS_SingleGeom: (index, factor, parameters)."""
    return [(LAMBERTIAN_IDX, fiso, (0.0, 0.0, 0.0)),
            (ROSSTHICK_IDX, fvol, (0.0, 0.0, 0.0)),
            (LISPARSE_IDX, fgeo, (float(param[0]), float(param[1]), 0.0))]


def coxmunk_setup(u10m: float, v10m: float, m: float):
    """Kernel list of VLIDORT_CoxMunk_SingleGeom (V283/VLIDORT_SurfaceMod.F90:
    815-1042): one Cox-Munk kernel of factor 1 with sigma^2 = 0.003 + 0.00512 W
    (default-real literals, W = |(u10m, v10m)| capped at 50 m/s), m^2, shadowing
    on (the supplement sets the third parameter to 1 for Cox-Munk kernels);
    100 azimuth streams."""
    w10 = math.sqrt(u10m * u10m + v10m * v10m)
    if w10 < 0:
        w10 = 0.0
    if w10 > 50:
        w10 = 50.0
    sigma2 = float(np.float32(0.003)) + float(np.float32(0.00512)) * w10
    return [(COXMUNK_IDX, float(np.float32(1.0)), (sigma2, m * m, 1.0))]


def vbrdf_master(nstreams: int, sza: float, vza: float, raa: float, kernels, nstreams_brdf_in: int = 50,
                 wsa: bool = True):
    """BRDF supplement outputs for one observation geometry.

    Returns a dict with ``db`` (exact direct-bounce BRDF), ``F[M,i,j]``
    (quadrature-quadrature, incident j -> reflected i), ``F0[M,i]``
    (solar -> quadrature i), ``UF[M,j]`` (quadrature j -> user), ``UF0[M]``
    (solar -> user) for M = 0 .. 2*nstreams-1, ``wsa`` per kernel,
    ``total_wsa`` and ``ok`` (False where the white-sky albedo check fails:
    the Fourier terms of the last kernel are then left out, as the reference
    leaves VBRDF_MAINMASTER at that point)."""
    nmom = 2 * nstreams - 1
    nbrdf = nstreams_brdf_in // 2
    qs, _ = getquad2(0.0, 1.0, nstreams)
    qsin = [math.sqrt(1.0 - x * x) for x in qs]
    mu0 = math.cos(sza * DEG_TO_RAD)
    s0 = math.sqrt(1.0 - mu0 * mu0)
    muv = math.cos(vza * DEG_TO_RAD)
    sv = math.sqrt(1.0 - muv * muv)
    phi = raa * DEG_TO_RAD
    cphi, sphi = math.cos(phi), math.sin(phi)
    xb, ab = getquad2(0.0, 1.0, nbrdf)
    xb = [PIE * x for x in xb]
    cxb = [math.cos(x) for x in xb]
    sxb = [math.sin(x) for x in xb]
    sq, sw = getquad2(0.0, 1.0, MAXSTREAMS_SCALING)
    ssin = [math.sqrt(1.0 - x * x) for x in sq]
    ssw = [sq[i] * sw[i] for i in range(MAXSTREAMS_SCALING)]

    F = np.zeros((nmom + 1, nstreams, nstreams))
    F0 = np.zeros((nmom + 1, nstreams))
    UF = np.zeros((nmom + 1, nstreams))
    UF0 = np.zeros(nmom + 1)
    db = 0.0
    wsa_k = []
    total_wsa = 0.0
    ok = True
    nk = len(kernels)
    for kk, (idx, factor, pars) in enumerate(kernels):
        fn = KERNELS[idx]
        lamb = idx == LAMBERTIAN_IDX
        db = db + factor * fn(pars, mu0, s0, muv, sv, phi, cphi, sphi)
        if wsa:
            w = 1.0
            if not lamb:
                help_a = 0.0
                for i in range(MAXSTREAMS_SCALING):
                    row = [_dot([fn(pars, sq[j], ssin[j], sq[i], ssin[i], xb[k], cxb[k], sxb[k])
                                 for k in range(nbrdf)], ab) for j in range(MAXSTREAMS_SCALING)]
                    help_a = help_a + _dot(row, ssw) * ssw[i]
                w = help_a * 4.0 / PIE
            wsa_k.append(w)
            total_wsa = total_wsa + factor * w
            if kk == nk - 1 and (total_wsa <= 0.0 or total_wsa > 1.0):
                ok = False
                break
        if lamb:
            F[0] += factor * 1.0
            F0[0] += factor * 1.0
            UF[0] += factor * 1.0
            UF0[0] += factor * 1.0
            continue
        k_ij = [[[fn(pars, qs[j], qsin[j], qs[i], qsin[i], xb[k], cxb[k], sxb[k]) for k in range(nbrdf)]
                 for j in range(nstreams)] for i in range(nstreams)]
        k_0i = [[fn(pars, mu0, s0, qs[i], qsin[i], xb[k], cxb[k], sxb[k]) for k in range(nbrdf)]
                for i in range(nstreams)]
        k_uj = [[fn(pars, qs[j], qsin[j], muv, sv, xb[k], cxb[k], sxb[k]) for k in range(nbrdf)]
                for j in range(nstreams)]
        k_u0 = [fn(pars, mu0, s0, muv, sv, xb[k], cxb[k], sxb[k]) for k in range(nbrdf)]
        for m in range(nmom + 1):
            if m == 0:
                delfac, caz = 1.0, list(ab)
            else:
                delfac = 2.0
                caz = [ab[k] * math.cos(float(m) * xb[k]) for k in range(nbrdf)]
            for i in range(nstreams):
                for j in range(nstreams):
                    F[m, i, j] = F[m, i, j] + factor * (delfac * _dot(k_ij[i][j], caz))
                F0[m, i] = F0[m, i] + factor * (delfac * _dot(k_0i[i], caz))
            for j in range(nstreams):
                UF[m, j] = UF[m, j] + factor * (delfac * _dot(k_uj[j], caz))
            UF0[m] = UF0[m] + factor * (delfac * _dot(k_u0, caz))
    return dict(db=db, F=F, F0=F0, UF=UF, UF0=UF0, wsa=wsa_k, total_wsa=total_wsa, ok=ok)
