/*
 * specfem_oracle.c -- CPU restatement of SPECFEM3D Cartesian's single-partition
 * solver-loop kernels, used ONLY as a test oracle and as the CPU baseline timed
 * beside the GPU path. Nothing in the product links or calls this file.
 *
 * Compiled twice: -DORC_REAL=float (FP32 oracle, same arithmetic order as the
 * reference CUSTOM_REAL=4 build) and -DORC_REAL=double (FP64 oracle). Build with
 * -ffp-contract=off so that no multiply-add is fused, as in the reference's
 * x86-64 flang build.
 *
 * Layouts (0-based):
 *   ibool[e*125 + p], p = i + 5 j + 25 k (Fortran ibool(i,j,k,ispec) order)
 *   vector fields [3*iglob + c] (Fortran displ(c,iglob))
 *   per-local-point arrays [e*125 + p]; attenuation memory [(e*125 + p)*3 + sls]
 *   hprime[i*5 + l] = derivative of Lagrange polynomial l at GLL node i
 *   hprimewgll[i*5 + l] = hprime[i*5+l] * w_i ; wgllwgll[a*5 + b] = w_a w_b
 *   jac9[(irr*9 + t)*125 + p], t = xix,xiy,xiz,etax,etay,etaz,gammax,gammay,gammaz
 *
 * Each function cites the reference routine it restates.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#ifndef ORC_REAL
#define ORC_REAL float
#endif
typedef ORC_REAL R;

#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#ifndef ORC_SUFFIX
#define ORC_SUFFIX _f32
#endif
#define FN(name) CAT(name, ORC_SUFFIX)

#define NGLL 5
#define NGLL3 125

/* -DORC_OMP builds the multi-threaded CPU baseline (liboracle_omp.so): element loops run in
 * parallel with atomic scatter-adds, so the summation order at shared points varies between
 * runs. The parity tests use the serial build only. */
#ifdef ORC_OMP
#include <omp.h>
#define ORC_PRAGMA(x) _Pragma(#x)
#define ORC_PAR_FOR ORC_PRAGMA(omp parallel for schedule(static))
#define ORC_ATOMIC ORC_PRAGMA(omp atomic)
void FN(orc_set_threads)(int n) { omp_set_num_threads(n); }
#else
#define ORC_PAR_FOR
#define ORC_ATOMIC
#endif

/* ------------------------------------------------------------------------- */
/* Newmark predictor: update_displ_elastic, update_displacement_scheme.f90:359-410
 * (and update_displ_acoustic :206-227 with n scalar values). */
void FN(orc_update_displ)(int64_t n, R *displ, R *veloc, R *accel, R deltat, R deltatover2, R deltatsqover2)
{
    ORC_PAR_FOR
    for (int64_t i = 0; i < n; i++) {
        displ[i] = displ[i] + deltat * veloc[i] + deltatsqover2 * accel[i];
        veloc[i] = veloc[i] + deltatover2 * accel[i];
        accel[i] = (R)0;
    }
}

/* mass division + corrector: compute_forces_viscoelastic_calling_routine.F90:368-374
 * followed by update_veloc_elastic (update_displacement_scheme.f90:646-693). */
void FN(orc_kernel3_elastic)(int64_t nglob, R *veloc, R *accel, const R *rmassx, const R *rmassy,
                             const R *rmassz, R deltatover2)
{
    ORC_PAR_FOR
    for (int64_t i = 0; i < nglob; i++) {
        accel[3 * i + 0] = accel[3 * i + 0] * rmassx[i];
        accel[3 * i + 1] = accel[3 * i + 1] * rmassy[i];
        accel[3 * i + 2] = accel[3 * i + 2] * rmassz[i];
    }
    ORC_PAR_FOR
    for (int64_t i = 0; i < 3 * nglob; i++) veloc[i] = veloc[i] + deltatover2 * accel[i];
}

/* acoustic mass + corrector: compute_forces_acoustic_calling_routine.f90 (rmass_acoustic, then
 * update_potential_dot_acoustic). */
void FN(orc_kernel3_acoustic)(int64_t nglob, R *potd, R *potdd, const R *rmass_ac, R deltatover2)
{
    for (int64_t i = 0; i < nglob; i++) potdd[i] = potdd[i] * rmass_ac[i];
    for (int64_t i = 0; i < nglob; i++) potd[i] = potd[i] + deltatover2 * potdd[i];
}

/* ------------------------------------------------------------------------- */
static inline R sum5_hA(const R *h, int row, const R *v, int s0, int stride)
{
    /* sum_l h[row*5+l] * v[s0 + l*stride], l = 0..4 in order (mxm5 rows) */
    R a = h[row * 5 + 0] * v[s0];
    a = a + h[row * 5 + 1] * v[s0 + stride];
    a = a + h[row * 5 + 2] * v[s0 + 2 * stride];
    a = a + h[row * 5 + 3] * v[s0 + 3 * stride];
    a = a + h[row * 5 + 4] * v[s0 + 4 * stride];
    return a;
}

static inline R sum5_hT(const R *h, int col, const R *v, int s0, int stride)
{
    /* sum_l h[l*5+col] * v[s0 + l*stride] */
    R a = h[0 * 5 + col] * v[s0];
    a = a + h[1 * 5 + col] * v[s0 + stride];
    a = a + h[2 * 5 + col] * v[s0 + 2 * stride];
    a = a + h[3 * 5 + col] * v[s0 + 3 * stride];
    a = a + h[4 * 5 + col] * v[s0 + 4 * stride];
    return a;
}

/* gradient in reference coordinates of a 125-point field:
 * t1 = sum_l hprime(i,l) f(l,j,k); t2 = sum_l hprime(j,l) f(i,l,k); t3 = sum_l hprime(k,l) f(i,j,l)
 * (compute_forces_viscoelastic.F90:331-335, mxm5_3comp_* ; same order of the five terms). */
static void grad_ref(const R *hp, const R *f, R *t1, R *t2, R *t3)
{
    for (int k = 0; k < 5; k++)
        for (int j = 0; j < 5; j++)
            for (int i = 0; i < 5; i++) {
                int p = i + 5 * j + 25 * k;
                t1[p] = sum5_hA(hp, i, f, 5 * j + 25 * k, 1);
                t2[p] = sum5_hA(hp, j, f, i + 25 * k, 5);
                t3[p] = sum5_hA(hp, k, f, i + 5 * j, 25);
            }
}

/* weighted divergence: n1 = sum_l hprimewgll(l,i) t1(l,j,k), etc. (:722-726) */
static void div_ref(const R *hpw, const R *t1, const R *t2, const R *t3, R *n1, R *n2, R *n3)
{
    for (int k = 0; k < 5; k++)
        for (int j = 0; j < 5; j++)
            for (int i = 0; i < 5; i++) {
                int p = i + 5 * j + 25 * k;
                n1[p] = sum5_hT(hpw, i, t1, 5 * j + 25 * k, 1);
                n2[p] = sum5_hT(hpw, j, t2, i + 25 * k, 5);
                n3[p] = sum5_hT(hpw, k, t3, i + 5 * j, 25);
            }
}

/* ------------------------------------------------------------------------- */
/* Elastic / viscoelastic element forces: compute_forces_viscoelastic.F90:33-962 (isotropic,
 * no PML/LDDRK/LTS), attenuation memory update compute_element_att_memory.f90:28-156.
 *   elems[nelem]: element list in processing order (phase_ispec_inner_elastic)
 *   att != 0: subtract the SLS memory stresses and update them (needs strain != 0)
 *   strain != 0: store deviatoric strain (and, if eps_trace_over_3 != NULL, trace/3)
 */
void FN(orc_forces_elastic)(int nelem, const int *elems, const int *ibool, const int *irregular,
                            R xix_regular, R jacobian_regular, const R *jac9, const R *jacobian,
                            const R *hprime, const R *hprimewgll, const R *wgllwgll,
                            const R *kappastore, const R *mustore,
                            const R *displ, const R *veloc, R *accel, R deltat,
                            int att, int strain,
                            const R *alphaval, const R *betaval, const R *gammaval,
                            const R *factor_common, const R *factor_common_kappa,
                            R *R_xx, R *R_yy, R *R_xy, R *R_xz, R *R_yz, R *R_trace,
                            R *eps_xx, R *eps_yy, R *eps_xy, R *eps_xz, R *eps_yz, R *eps_trace,
                            R *eps_trace_over_3)
{
    const R ONE_THIRD = (R)1 / (R)3;
    const R FOUR_THIRDS = (R)4 / (R)3;

    ORC_PAR_FOR
    for (int ie = 0; ie < nelem; ie++) {
        R dx[125], dy[125], dz[125];
        R tx1[125], tx2[125], tx3[125], ty1[125], ty2[125], ty3[125], tz1[125], tz2[125], tz3[125];
        R ax[125], ay[125], az[125];
        R ax1[125], ax2[125], ax3[125], ay1[125], ay2[125], ay3[125], az1[125], az2[125], az3[125];
        R nx1[125], nx2[125], nx3[125], ny1[125], ny2[125], ny3[125], nz1[125], nz2[125], nz3[125];
        R e_tr[125], e_xx[125], e_yy[125], e_xy[125], e_xz[125], e_yz[125];
        const int e = elems[ie];
        const int *ib = ibool + (int64_t)e * 125;
        for (int p = 0; p < 125; p++) {
            int64_t g = ib[p];
            dx[p] = displ[3 * g + 0];
            dy[p] = displ[3 * g + 1];
            dz[p] = displ[3 * g + 2];
        }
        grad_ref(hprime, dx, tx1, tx2, tx3);
        grad_ref(hprime, dy, ty1, ty2, ty3);
        grad_ref(hprime, dz, tz1, tz2, tz3);

        const int irr = irregular[e];
        const R *J = irr ? jac9 + (int64_t)(irr - 1) * 9 * 125 : 0;

        if (strain && att) {
            /* strain at u + dt*v (compute_strain_in_element_att): temp_att = temp + sum h * dt*v */
            for (int p = 0; p < 125; p++) {
                int64_t g = ib[p];
                ax[p] = deltat * veloc[3 * g + 0];
                ay[p] = deltat * veloc[3 * g + 1];
                az[p] = deltat * veloc[3 * g + 2];
            }
            for (int k = 0; k < 5; k++)
                for (int j = 0; j < 5; j++)
                    for (int i = 0; i < 5; i++) {
                        int p = i + 5 * j + 25 * k;
                        R a1 = tx1[p], a2 = tx2[p], a3 = tx3[p], b1 = ty1[p], b2 = ty2[p], b3 = ty3[p];
                        R c1 = tz1[p], c2 = tz2[p], c3 = tz3[p];
                        for (int l = 0; l < 5; l++) {
                            R hp1 = hprime[i * 5 + l], hp2 = hprime[j * 5 + l], hp3 = hprime[k * 5 + l];
                            int q1 = l + 5 * j + 25 * k, q2 = i + 5 * l + 25 * k, q3 = i + 5 * j + 25 * l;
                            a1 = a1 + ax[q1] * hp1; b1 = b1 + ay[q1] * hp1; c1 = c1 + az[q1] * hp1;
                            a2 = a2 + ax[q2] * hp2; b2 = b2 + ay[q2] * hp2; c2 = c2 + az[q2] * hp2;
                            a3 = a3 + ax[q3] * hp3; b3 = b3 + ay[q3] * hp3; c3 = c3 + az[q3] * hp3;
                        }
                        ax1[p] = a1; ax2[p] = a2; ax3[p] = a3;
                        ay1[p] = b1; ay2[p] = b2; ay3[p] = b3;
                        az1[p] = c1; az2[p] = c2; az3[p] = c3;
                    }
        }

        for (int p = 0; p < 125; p++) {
            R duxdx, duxdy, duxdz, duydx, duydy, duydz, duzdx, duzdy, duzdz;
            R xixl = 0, xiyl = 0, xizl = 0, etaxl = 0, etayl = 0, etazl = 0, gaxl = 0, gayl = 0, gazl = 0;
            if (irr) {
                xixl = J[0 * 125 + p]; xiyl = J[1 * 125 + p]; xizl = J[2 * 125 + p];
                etaxl = J[3 * 125 + p]; etayl = J[4 * 125 + p]; etazl = J[5 * 125 + p];
                gaxl = J[6 * 125 + p]; gayl = J[7 * 125 + p]; gazl = J[8 * 125 + p];
                duxdx = xixl * tx1[p] + etaxl * tx2[p] + gaxl * tx3[p];
                duxdy = xiyl * tx1[p] + etayl * tx2[p] + gayl * tx3[p];
                duxdz = xizl * tx1[p] + etazl * tx2[p] + gazl * tx3[p];
                duydx = xixl * ty1[p] + etaxl * ty2[p] + gaxl * ty3[p];
                duydy = xiyl * ty1[p] + etayl * ty2[p] + gayl * ty3[p];
                duydz = xizl * ty1[p] + etazl * ty2[p] + gazl * ty3[p];
                duzdx = xixl * tz1[p] + etaxl * tz2[p] + gaxl * tz3[p];
                duzdy = xiyl * tz1[p] + etayl * tz2[p] + gayl * tz3[p];
                duzdz = xizl * tz1[p] + etazl * tz2[p] + gazl * tz3[p];
            } else {
                duxdx = xix_regular * tx1[p]; duxdy = xix_regular * tx2[p]; duxdz = xix_regular * tx3[p];
                duydx = xix_regular * ty1[p]; duydy = xix_regular * ty2[p]; duydz = xix_regular * ty3[p];
                duzdx = xix_regular * tz1[p]; duzdy = xix_regular * tz2[p]; duzdz = xix_regular * tz3[p];
            }

            if (strain) {
                R sxx, syy, szz, sxy, syx, sxz, szx, syz, szy;
                if (att) {
                    if (irr) {
                        sxx = xixl * ax1[p] + etaxl * ax2[p] + gaxl * ax3[p];
                        sxy = xiyl * ax1[p] + etayl * ax2[p] + gayl * ax3[p];
                        sxz = xizl * ax1[p] + etazl * ax2[p] + gazl * ax3[p];
                        syx = xixl * ay1[p] + etaxl * ay2[p] + gaxl * ay3[p];
                        syy = xiyl * ay1[p] + etayl * ay2[p] + gayl * ay3[p];
                        syz = xizl * ay1[p] + etazl * ay2[p] + gazl * ay3[p];
                        szx = xixl * az1[p] + etaxl * az2[p] + gaxl * az3[p];
                        szy = xiyl * az1[p] + etayl * az2[p] + gayl * az3[p];
                        szz = xizl * az1[p] + etazl * az2[p] + gazl * az3[p];
                    } else {
                        sxx = xix_regular * ax1[p]; sxy = xix_regular * ax2[p]; sxz = xix_regular * ax3[p];
                        syx = xix_regular * ay1[p]; syy = xix_regular * ay2[p]; syz = xix_regular * ay3[p];
                        szx = xix_regular * az1[p]; szy = xix_regular * az2[p]; szz = xix_regular * az3[p];
                    }
                } else {
                    sxx = duxdx; sxy = duxdy; sxz = duxdz; syx = duydx; syy = duydy; syz = duydz;
                    szx = duzdx; szy = duzdy; szz = duzdz;
                }
                /* here s<c><d> = d u_c / d x_d */
                R xy_p = sxy + syx, zx_p = szx + sxz, zy_p = szy + syz;
                e_tr[p] = (sxx + syy + szz);
                R templ = ONE_THIRD * (sxx + syy + szz);
                if (eps_trace_over_3) eps_trace_over_3[(int64_t)e * 125 + p] = templ;
                e_xx[p] = sxx - templ;
                e_yy[p] = syy - templ;
                e_xy[p] = (R)0.5 * xy_p;
                e_xz[p] = (R)0.5 * zx_p;
                e_yz[p] = (R)0.5 * zy_p;
            }

            R duxdyl_plus_duydxl = duxdy + duydx;
            R duzdxl_plus_duxdzl = duzdx + duxdz;
            R duzdyl_plus_duydzl = duzdy + duydz;
            R kappal = kappastore[(int64_t)e * 125 + p];
            R mul = mustore[(int64_t)e * 125 + p];
            R lp2m = kappal + FOUR_THIRDS * mul;
            R lambdal = lp2m - (R)2 * mul;
            R sigma_xx = lp2m * duxdx + lambdal * (duydy + duzdz);
            R sigma_yy = lp2m * duydy + lambdal * (duxdx + duzdz);
            R sigma_zz = lp2m * duzdz + lambdal * (duxdx + duydy);
            R sigma_xy = mul * duxdyl_plus_duydxl;
            R sigma_xz = mul * duzdxl_plus_duxdzl;
            R sigma_yz = mul * duzdyl_plus_duydzl;

            if (att) {
                int64_t q = ((int64_t)e * 125 + p) * 3;
                R rxx = (R_xx[q] + R_xx[q + 1]) + R_xx[q + 2];
                R ryy = (R_yy[q] + R_yy[q + 1]) + R_yy[q + 2];
                R rtr = (R_trace[q] + R_trace[q + 1]) + R_trace[q + 2];
                sigma_xx = sigma_xx - rxx - rtr;
                sigma_yy = sigma_yy - ryy - rtr;
                sigma_zz = sigma_zz + rxx + ryy - rtr;
                sigma_xy = sigma_xy - ((R_xy[q] + R_xy[q + 1]) + R_xy[q + 2]);
                sigma_xz = sigma_xz - ((R_xz[q] + R_xz[q + 1]) + R_xz[q + 2]);
                sigma_yz = sigma_yz - ((R_yz[q] + R_yz[q + 1]) + R_yz[q + 2]);
            }
            R sigma_yx = sigma_xy, sigma_zx = sigma_xz, sigma_zy = sigma_yz;
            if (irr) {
                R jl = jacobian[(int64_t)(irr - 1) * 125 + p];
                tx1[p] = jl * (sigma_xx * xixl + sigma_yx * xiyl + sigma_zx * xizl);
                ty1[p] = jl * (sigma_xy * xixl + sigma_yy * xiyl + sigma_zy * xizl);
                tz1[p] = jl * (sigma_xz * xixl + sigma_yz * xiyl + sigma_zz * xizl);
                tx2[p] = jl * (sigma_xx * etaxl + sigma_yx * etayl + sigma_zx * etazl);
                ty2[p] = jl * (sigma_xy * etaxl + sigma_yy * etayl + sigma_zy * etazl);
                tz2[p] = jl * (sigma_xz * etaxl + sigma_yz * etayl + sigma_zz * etazl);
                tx3[p] = jl * (sigma_xx * gaxl + sigma_yx * gayl + sigma_zx * gazl);
                ty3[p] = jl * (sigma_xy * gaxl + sigma_yy * gayl + sigma_zy * gazl);
                tz3[p] = jl * (sigma_xz * gaxl + sigma_yz * gayl + sigma_zz * gazl);
            } else {
                R jl = jacobian_regular;
                tx1[p] = jl * sigma_xx * xix_regular;
                ty1[p] = jl * sigma_xy * xix_regular;
                tz1[p] = jl * sigma_xz * xix_regular;
                tx2[p] = jl * sigma_yx * xix_regular;
                ty2[p] = jl * sigma_yy * xix_regular;
                tz2[p] = jl * sigma_yz * xix_regular;
                tx3[p] = jl * sigma_zx * xix_regular;
                ty3[p] = jl * sigma_zy * xix_regular;
                tz3[p] = jl * sigma_zz * xix_regular;
            }
        }

        div_ref(hprimewgll, tx1, tx2, tx3, nx1, nx2, nx3);
        div_ref(hprimewgll, ty1, ty2, ty3, ny1, ny2, ny3);
        div_ref(hprimewgll, tz1, tz2, tz3, nz1, nz2, nz3);

        for (int k = 0; k < 5; k++)
            for (int j = 0; j < 5; j++)
                for (int i = 0; i < 5; i++) {
                    int p = i + 5 * j + 25 * k;
                    int64_t g = ib[p];
                    R fac1 = wgllwgll[j * 5 + k];
                    R fac2 = wgllwgll[i * 5 + k];
                    R fac3 = wgllwgll[i * 5 + j];
                    const R fx = fac1 * nx1[p] + fac2 * nx2[p] + fac3 * nx3[p];
                    const R fy = fac1 * ny1[p] + fac2 * ny2[p] + fac3 * ny3[p];
                    const R fz = fac1 * nz1[p] + fac2 * nz2[p] + fac3 * nz3[p];
                    ORC_ATOMIC
                    accel[3 * g + 0] -= fx;
                    ORC_ATOMIC
                    accel[3 * g + 1] -= fy;
                    ORC_ATOMIC
                    accel[3 * g + 2] -= fz;
                }

        if (att) {
            /* compute_element_att_memory_second_order_rk */
            for (int p = 0; p < 125; p++) {
                int64_t ep = (int64_t)e * 125 + p;
                int64_t q = ep * 3;
                for (int s = 0; s < 3; s++) {
                    R fl = kappastore[ep] * factor_common_kappa[q + s];
                    R Sn = eps_trace[ep], Snp1 = e_tr[p];
                    R_trace[q + s] = alphaval[s] * R_trace[q + s] + fl * (betaval[s] * Sn + gammaval[s] * Snp1);
                }
                for (int s = 0; s < 3; s++) {
                    R fl = mustore[ep] * factor_common[q + s];
                    R_xx[q + s] = alphaval[s] * R_xx[q + s] + fl * (betaval[s] * eps_xx[ep] + gammaval[s] * e_xx[p]);
                    R_yy[q + s] = alphaval[s] * R_yy[q + s] + fl * (betaval[s] * eps_yy[ep] + gammaval[s] * e_yy[p]);
                    R_xy[q + s] = alphaval[s] * R_xy[q + s] + fl * (betaval[s] * eps_xy[ep] + gammaval[s] * e_xy[p]);
                    R_xz[q + s] = alphaval[s] * R_xz[q + s] + fl * (betaval[s] * eps_xz[ep] + gammaval[s] * e_xz[p]);
                    R_yz[q + s] = alphaval[s] * R_yz[q + s] + fl * (betaval[s] * eps_yz[ep] + gammaval[s] * e_yz[p]);
                }
            }
        }
        if (strain) {
            int64_t b = (int64_t)e * 125;
            if (att) memcpy(eps_trace + b, e_tr, sizeof(e_tr));
            memcpy(eps_xx + b, e_xx, sizeof(e_xx));
            memcpy(eps_yy + b, e_yy, sizeof(e_yy));
            memcpy(eps_xy + b, e_xy, sizeof(e_xy));
            memcpy(eps_xz + b, e_xz, sizeof(e_xz));
            memcpy(eps_yz + b, e_yz, sizeof(e_yz));
        }
    }
}

/* Acoustic element forces: compute_forces_acoustic.F90:35-726 (no PML). */
void FN(orc_forces_acoustic)(int nelem, const int *elems, const int *ibool, const int *irregular,
                             R xix_regular, R jacobian_regular, const R *jac9, const R *jacobian,
                             const R *hprime, const R *hprimewgll, const R *wgllwgll,
                             const R *rhostore, const R *pot, R *potdd)
{
    ORC_PAR_FOR
    for (int ie = 0; ie < nelem; ie++) {
        R chi[125], t1[125], t2[125], t3[125], n1[125], n2[125], n3[125];
        const int e = elems[ie];
        const int *ib = ibool + (int64_t)e * 125;
        for (int p = 0; p < 125; p++) chi[p] = pot[ib[p]];
        grad_ref(hprime, chi, t1, t2, t3);
        const int irr = irregular[e];
        if (irr) {
            const R *J = jac9 + (int64_t)(irr - 1) * 9 * 125;
            for (int p = 0; p < 125; p++) {
                R rho_invl = (R)1 / rhostore[(int64_t)e * 125 + p];
                R xixl = J[0 * 125 + p], xiyl = J[1 * 125 + p], xizl = J[2 * 125 + p];
                R etaxl = J[3 * 125 + p], etayl = J[4 * 125 + p], etazl = J[5 * 125 + p];
                R gaxl = J[6 * 125 + p], gayl = J[7 * 125 + p], gazl = J[8 * 125 + p];
                R jl = jacobian[(int64_t)(irr - 1) * 125 + p];
                R dpdx = xixl * t1[p] + etaxl * t2[p] + gaxl * t3[p];
                R dpdy = xiyl * t1[p] + etayl * t2[p] + gayl * t3[p];
                R dpdz = xizl * t1[p] + etazl * t2[p] + gazl * t3[p];
                t1[p] = rho_invl * jl * (xixl * dpdx + xiyl * dpdy + xizl * dpdz);
                t2[p] = rho_invl * jl * (etaxl * dpdx + etayl * dpdy + etazl * dpdz);
                t3[p] = rho_invl * jl * (gaxl * dpdx + gayl * dpdy + gazl * dpdz);
            }
        } else {
            R jl = jacobian_regular;
            for (int p = 0; p < 125; p++) {
                R rho_invl = (R)1 / rhostore[(int64_t)e * 125 + p];
                t1[p] = rho_invl * jl * xix_regular * xix_regular * t1[p];
                t2[p] = rho_invl * jl * xix_regular * xix_regular * t2[p];
                t3[p] = rho_invl * jl * xix_regular * xix_regular * t3[p];
            }
        }
        div_ref(hprimewgll, t1, t2, t3, n1, n2, n3);
        for (int k = 0; k < 5; k++)
            for (int j = 0; j < 5; j++)
                for (int i = 0; i < 5; i++) {
                    int p = i + 5 * j + 25 * k;
                    const R f = wgllwgll[j * 5 + k] * n1[p] + wgllwgll[i * 5 + k] * n2[p] + wgllwgll[i * 5 + j] * n3[p];
                    ORC_ATOMIC
                    potdd[ib[p]] -= f;
                }
    }
}

/* ------------------------------------------------------------------------- */
/* Stacey elastic: compute_stacey_viscoelastic.F90:32-202. b_absorb (nullable) receives the
 * traction terms per face point [(f*25 + igll)*3 + c] when storing for a later replay. */
void FN(orc_stacey_elastic)(int nfaces, const int *face_ispec, const int *face_ijk, const R *normal,
                            const R *jacw, const int *ibool, const unsigned char *is_elastic,
                            const R *rho_vp, const R *rho_vs, const R *veloc, R *accel, R *b_absorb)
{
    for (int f = 0; f < nfaces; f++) {
        int e = face_ispec[f];
        if (!is_elastic[e]) continue;
        for (int a = 0; a < 25; a++) {
            const int *ijk = face_ijk + ((int64_t)f * 25 + a) * 3;
            int p = ijk[0] + 5 * ijk[1] + 25 * ijk[2];
            int64_t g = ibool[(int64_t)e * 125 + p];
            R vx = veloc[3 * g], vy = veloc[3 * g + 1], vz = veloc[3 * g + 2];
            const R *n = normal + ((int64_t)f * 25 + a) * 3;
            R nx = n[0], ny = n[1], nz = n[2];
            R vn = vx * nx + vy * ny + vz * nz;
            R rvp = rho_vp[(int64_t)e * 125 + p], rvs = rho_vs[(int64_t)e * 125 + p];
            R tx = rvp * vn * nx + rvs * (vx - vn * nx);
            R ty = rvp * vn * ny + rvs * (vy - vn * ny);
            R tz = rvp * vn * nz + rvs * (vz - vn * nz);
            R jw = jacw[(int64_t)f * 25 + a];
            accel[3 * g] = accel[3 * g] - tx * jw;
            accel[3 * g + 1] = accel[3 * g + 1] - ty * jw;
            accel[3 * g + 2] = accel[3 * g + 2] - tz * jw;
            if (b_absorb) {
                b_absorb[((int64_t)f * 25 + a) * 3 + 0] = tx * jw;
                b_absorb[((int64_t)f * 25 + a) * 3 + 1] = ty * jw;
                b_absorb[((int64_t)f * 25 + a) * 3 + 2] = tz * jw;
            }
        }
    }
}

/* Stacey acoustic: compute_stacey_acoustic.f90:30-120. */
void FN(orc_stacey_acoustic)(int nfaces, const int *face_ispec, const int *face_ijk, const R *jacw,
                             const int *ibool, const unsigned char *is_acoustic, const R *rhostore,
                             const R *kappastore, const R *potd, R *potdd)
{
    for (int f = 0; f < nfaces; f++) {
        int e = face_ispec[f];
        if (!is_acoustic[e]) continue;
        for (int a = 0; a < 25; a++) {
            const int *ijk = face_ijk + ((int64_t)f * 25 + a) * 3;
            int p = ijk[0] + 5 * ijk[1] + 25 * ijk[2];
            int64_t g = ibool[(int64_t)e * 125 + p];
            R rhol = rhostore[(int64_t)e * 125 + p];
            R cpl = (R)sqrt(kappastore[(int64_t)e * 125 + p] / rhol);
            R jw = jacw[(int64_t)f * 25 + a];
            R absorbl = potd[g] * jw / cpl / rhol;
            potdd[g] = potdd[g] - absorbl;
        }
    }
}

/* fluid-solid coupling, acoustic side: compute_coupling_acoustic_el.f90:30-158 (sign = -1 for
 * the adjoint wavefield of a kernel run). */
void FN(orc_coupling_ac_el)(int nfaces, const int *face_ispec, const int *face_ijk, const R *normal,
                            const R *jacw, const int *ibool, const R *displ, R *potdd, int flip)
{
    for (int f = 0; f < nfaces; f++) {
        int e = face_ispec[f];
        for (int a = 0; a < 25; a++) {
            const int *ijk = face_ijk + ((int64_t)f * 25 + a) * 3;
            int p = ijk[0] + 5 * ijk[1] + 25 * ijk[2];
            int64_t g = ibool[(int64_t)e * 125 + p];
            R ux = displ[3 * g], uy = displ[3 * g + 1], uz = displ[3 * g + 2];
            if (flip) { ux = -ux; uy = -uy; uz = -uz; }
            const R *n = normal + ((int64_t)f * 25 + a) * 3;
            R un = ux * n[0] + uy * n[1] + uz * n[2];
            potdd[g] = potdd[g] + jacw[(int64_t)f * 25 + a] * un;
        }
    }
}

/* fluid-solid coupling, elastic side: compute_coupling_viscoelastic_ac.f90:30-165. */
void FN(orc_coupling_el_ac)(int nfaces, const int *face_ispec, const int *face_ijk, const R *normal,
                            const R *jacw, const int *ibool, const R *potdd, R *accel, int flip)
{
    for (int f = 0; f < nfaces; f++) {
        int e = face_ispec[f];
        for (int a = 0; a < 25; a++) {
            const int *ijk = face_ijk + ((int64_t)f * 25 + a) * 3;
            int p = ijk[0] + 5 * ijk[1] + 25 * ijk[2];
            int64_t g = ibool[(int64_t)e * 125 + p];
            R pr = -potdd[g];
            if (flip) pr = -pr;
            const R *n = normal + ((int64_t)f * 25 + a) * 3;
            R jw = jacw[(int64_t)f * 25 + a];
            accel[3 * g] = accel[3 * g] + jw * n[0] * pr;
            accel[3 * g + 1] = accel[3 * g + 1] + jw * n[1] * pr;
            accel[3 * g + 2] = accel[3 * g + 2] + jw * n[2] * pr;
        }
    }
}

/* acoustic free surface: compute_forces_acoustic_calling_routine.f90:754-808 */
void FN(orc_free_surface)(int nfaces, const int *face_ispec, const int *face_ijk, const int *ibool,
                          const unsigned char *is_acoustic, R *pot, R *potd, R *potdd)
{
    for (int f = 0; f < nfaces; f++) {
        int e = face_ispec[f];
        if (!is_acoustic[e]) continue;
        for (int a = 0; a < 25; a++) {
            const int *ijk = face_ijk + ((int64_t)f * 25 + a) * 3;
            int p = ijk[0] + 5 * ijk[1] + 25 * ijk[2];
            int64_t g = ibool[(int64_t)e * 125 + p];
            pot[g] = 0; potd[g] = 0; potdd[g] = 0;
        }
    }
}

/* point sources: compute_add_sources_viscoelastic.F90:120-160; sourcearrays[(s*125 + p)*3 + c];
 * stf given per source (already cast to CUSTOM_REAL by the caller, as stf_used). */
void FN(orc_add_sources_el)(int nsrc, const int *src_ispec, const R *sourcearrays, const R *stf,
                            const int *ibool, const unsigned char *is_elastic, R *accel)
{
    for (int s = 0; s < nsrc; s++) {
        int e = src_ispec[s];
        if (!is_elastic[e]) continue;
        R st = stf[s];
        for (int p = 0; p < 125; p++) {
            int64_t g = ibool[(int64_t)e * 125 + p];
            const R *sa = sourcearrays + ((int64_t)s * 125 + p) * 3;
            accel[3 * g] = accel[3 * g] + sa[0] * st;
            accel[3 * g + 1] = accel[3 * g + 1] + sa[1] * st;
            accel[3 * g + 2] = accel[3 * g + 2] + sa[2] * st;
        }
    }
}

/* point sources in acoustic elements: compute_add_sources_acoustic.f90:95-124,
 * potential_dot_dot -= sourcearrays(isource,1,i,j,k) * stf_used / kappastore(i,j,k,ispec). */
void FN(orc_add_sources_ac)(int nsrc, const int *src_ispec, const R *sourcearrays, const R *stf,
                            const int *ibool, const unsigned char *is_acoustic, const R *kappastore, R *potdd)
{
    for (int s = 0; s < nsrc; s++) {
        int e = src_ispec[s];
        if (!is_acoustic[e]) continue;
        R st = stf[s];
        for (int p = 0; p < 125; p++) {
            int64_t g = ibool[(int64_t)e * 125 + p];
            potdd[g] = potdd[g] - sourcearrays[((int64_t)s * 125 + p) * 3] * st / kappastore[(int64_t)e * 125 + p];
        }
    }
}

/* adjoint sources at receivers in acoustic elements: compute_add_sources_acoustic.f90:205-235,
 * potential_dot_dot += source_adjoint(1,irec,it_index) * hlagrange (CUSTOM_REAL interpolants). */
void FN(orc_add_adjoint_sources_ac)(int nrec, const int *rec_ispec, const double *hxir, const double *hetar,
                                    const double *hgammar, const R *adj, const int *ibool,
                                    const unsigned char *is_acoustic, R *potdd)
{
    for (int r = 0; r < nrec; r++) {
        int e = rec_ispec[r];
        if (!is_acoustic[e]) continue;
        for (int k = 0; k < 5; k++)
            for (int j = 0; j < 5; j++)
                for (int i = 0; i < 5; i++) {
                    int64_t g = ibool[(int64_t)e * 125 + i + 5 * j + 25 * k];
                    R h = (R)hxir[r * 5 + i] * (R)hetar[r * 5 + j] * (R)hgammar[r * 5 + k];
                    potdd[g] = potdd[g] + adj[r * 3] * h;
                }
    }
}

/* adjoint sources at receivers: compute_add_sources_viscoelastic.F90:226-245 (receivers in elastic
 * elements; is_elastic may be NULL when all are);
 * adj[r*3 + c] is the adjoint source value for this step, h* are the interpolants (double). */
void FN(orc_add_adjoint_sources_el)(int nrec, const int *rec_ispec, const double *hxir, const double *hetar,
                                    const double *hgammar, const R *adj, const int *ibool, R *accel,
                                    const unsigned char *is_elastic)
{
    for (int r = 0; r < nrec; r++) {
        int e = rec_ispec[r];
        if (is_elastic && !is_elastic[e]) continue;
        for (int k = 0; k < 5; k++)
            for (int j = 0; j < 5; j++)
                for (int i = 0; i < 5; i++) {
                    int64_t g = ibool[(int64_t)e * 125 + i + 5 * j + 25 * k];
                    /* hlagrange is CUSTOM_REAL: product of the stored CUSTOM_REAL interpolants */
                    R h = (R)hxir[r * 5 + i] * (R)hetar[r * 5 + j] * (R)hgammar[r * 5 + k];
                    accel[3 * g] = accel[3 * g] + adj[r * 3] * h;
                    accel[3 * g + 1] = accel[3 * g + 1] + adj[r * 3 + 1] * h;
                    accel[3 * g + 2] = accel[3 * g + 2] + adj[r * 3 + 2] * h;
                }
    }
}

/* receiver interpolation (elastic): compute_interpolated_dva_viscoelast, compute_interpolated_dva.f90
 * :28-100, accumulated in double; out[r*3 + c] rotated by nu[r*9 + 3*a + b]. */
void FN(orc_seismo_elastic)(int nrec, const int *rec_ispec, const double *hxir, const double *hetar,
                            const double *hgammar, const double *nu, const int *ibool, const R *field,
                            R *out, const unsigned char *is_elastic)
{
    for (int r = 0; r < nrec; r++) {
        int e = rec_ispec[r];
        if (is_elastic && !is_elastic[e]) continue;
        double dx = 0, dy = 0, dz = 0;
        for (int k = 0; k < 5; k++)
            for (int j = 0; j < 5; j++)
                for (int i = 0; i < 5; i++) {
                    int64_t g = ibool[(int64_t)e * 125 + i + 5 * j + 25 * k];
                    double h = hxir[r * 5 + i] * hetar[r * 5 + j] * hgammar[r * 5 + k];
                    dx = dx + (double)field[3 * g] * h;
                    dy = dy + (double)field[3 * g + 1] * h;
                    dz = dz + (double)field[3 * g + 2] * h;
                }
        const double *n = nu + (int64_t)r * 9;
        out[r * 3 + 0] = (R)(n[0] * dx + n[1] * dy + n[2] * dz);
        out[r * 3 + 1] = (R)(n[3] * dx + n[4] * dy + n[5] * dz);
        out[r * 3 + 2] = (R)(n[6] * dx + n[7] * dy + n[8] * dz);
    }
}

/* element strain from a displacement field: compute_element_strain.F90:28-130 */
void FN(orc_element_strain)(int nelem, const int *elems, const int *ibool, const int *irregular,
                            R xix_regular, const R *jac9, const R *hprime, const R *displ,
                            R *eps_xx, R *eps_yy, R *eps_xy, R *eps_xz, R *eps_yz, R *eps_trace_over_3)
{
    const R ONE_THIRD = (R)1 / (R)3;
    R dx[125], dy[125], dz[125];
    R tx1[125], tx2[125], tx3[125], ty1[125], ty2[125], ty3[125], tz1[125], tz2[125], tz3[125];
    for (int ie = 0; ie < nelem; ie++) {
        int e = elems[ie];
        const int *ib = ibool + (int64_t)e * 125;
        for (int p = 0; p < 125; p++) {
            int64_t g = ib[p];
            dx[p] = displ[3 * g]; dy[p] = displ[3 * g + 1]; dz[p] = displ[3 * g + 2];
        }
        /* the reference accumulates from 0: 0 + h1 f1 + ... equals our first-product start */
        grad_ref(hprime, dx, tx1, tx2, tx3);
        grad_ref(hprime, dy, ty1, ty2, ty3);
        grad_ref(hprime, dz, tz1, tz2, tz3);
        int irr = irregular[e];
        for (int p = 0; p < 125; p++) {
            R duxdx, duxdy, duxdz, duydx, duydy, duydz, duzdx, duzdy, duzdz;
            if (irr) {
                const R *J = jac9 + (int64_t)(irr - 1) * 9 * 125;
                R xixl = J[p], xiyl = J[125 + p], xizl = J[250 + p], etaxl = J[375 + p], etayl = J[500 + p];
                R etazl = J[625 + p], gaxl = J[750 + p], gayl = J[875 + p], gazl = J[1000 + p];
                duxdx = xixl * tx1[p] + etaxl * tx2[p] + gaxl * tx3[p];
                duxdy = xiyl * tx1[p] + etayl * tx2[p] + gayl * tx3[p];
                duxdz = xizl * tx1[p] + etazl * tx2[p] + gazl * tx3[p];
                duydx = xixl * ty1[p] + etaxl * ty2[p] + gaxl * ty3[p];
                duydy = xiyl * ty1[p] + etayl * ty2[p] + gayl * ty3[p];
                duydz = xizl * ty1[p] + etazl * ty2[p] + gazl * ty3[p];
                duzdx = xixl * tz1[p] + etaxl * tz2[p] + gaxl * tz3[p];
                duzdy = xiyl * tz1[p] + etayl * tz2[p] + gayl * tz3[p];
                duzdz = xizl * tz1[p] + etazl * tz2[p] + gazl * tz3[p];
            } else {
                duxdx = xix_regular * tx1[p]; duxdy = xix_regular * tx2[p]; duxdz = xix_regular * tx3[p];
                duydx = xix_regular * ty1[p]; duydy = xix_regular * ty2[p]; duydz = xix_regular * ty3[p];
                duzdx = xix_regular * tz1[p]; duzdy = xix_regular * tz2[p]; duzdz = xix_regular * tz3[p];
            }
            R templ = ONE_THIRD * (duxdx + duydy + duzdz);
            int64_t ep = (int64_t)e * 125 + p;
            eps_trace_over_3[ep] = templ;
            eps_xx[ep] = duxdx - templ;
            eps_yy[ep] = duydy - templ;
            eps_xy[ep] = (R)0.5 * (duxdy + duydx);
            eps_xz[ep] = (R)0.5 * (duzdx + duxdz);
            eps_yz[ep] = (R)0.5 * (duzdy + duydz);
        }
    }
}

/* sensitivity-kernel accumulation: compute_kernels.f90:67-222 (isotropic). */
void FN(orc_compute_kernels_el)(int nelem, const int *elems, const int *ibool, const R *accel,
                                const R *b_displ, const R *exx, const R *eyy, const R *exy, const R *exz,
                                const R *eyz, const R *etr3, const R *bxx, const R *byy, const R *bxy,
                                const R *bxz, const R *byz, const R *btr3, R deltat, R *rho_kl, R *mu_kl,
                                R *kappa_kl)
{
    for (int ie = 0; ie < nelem; ie++) {
        int e = elems[ie];
        for (int p = 0; p < 125; p++) {
            int64_t ep = (int64_t)e * 125 + p;
            int64_t g = ibool[ep];
            R dot = accel[3 * g] * b_displ[3 * g] + accel[3 * g + 1] * b_displ[3 * g + 1]
                    + accel[3 * g + 2] * b_displ[3 * g + 2];
            rho_kl[ep] = rho_kl[ep] + deltat * dot;
            R e1 = exx[ep], e2 = eyy[ep], e3 = exy[ep], e4 = exz[ep], e5 = eyz[ep];
            R b1 = bxx[ep], b2 = byy[ep], b3 = bxy[ep], b4 = bxz[ep], b5 = byz[ep];
            mu_kl[ep] = mu_kl[ep] + deltat * (e1 * b1 + e2 * b2 + (e1 + e2) * (b1 + b2)
                                              + (R)2 * (e3 * b3 + e4 * b4 + e5 * b5));
            kappa_kl[ep] = kappa_kl[ep] + deltat * ((R)9 * etr3[ep] * btr3[ep]);
        }
    }
}

/* backward Stacey of the time-reversed reconstruction: b_accel -= b_absorb_field, the record the
 * forward run stored at this step (compute_stacey_viscoelastic.F90:212-287, face order) */
void FN(orc_stacey_elastic_replay)(int nfaces, const int *face_ispec, const int *face_ijk, const int *ibool,
                                   const unsigned char *is_elastic, const R *b_absorb, R *accel)
{
    for (int f = 0; f < nfaces; f++) {
        int e = face_ispec[f];
        if (!is_elastic[e]) continue;
        for (int a = 0; a < 25; a++) {
            const int *ijk = face_ijk + ((int64_t)f * 25 + a) * 3;
            int p = ijk[0] + 5 * ijk[1] + 25 * ijk[2];
            int64_t g = ibool[(int64_t)e * 125 + p];
            for (int c = 0; c < 3; c++) accel[3 * g + c] = accel[3 * g + c] - b_absorb[((int64_t)f * 25 + a) * 3 + c];
        }
    }
}

/* acoustic sensitivity kernels: compute_kernels.f90 compute_kernels_ac (rho_ac_kl, kappa_ac_kl) with
 * the gradient of compute_gradient_in_acoustic_fast_Deville (compute_gradient_in_acoustic.f90):
 * grad chi / rho from the three contractions, mapped by the inverse Jacobian (irregular) or
 * xix_regular, then multiplied back by rho; rho_ac_kl += dt grad(chi_adj).grad(chi_b),
 * kappa_ac_kl += dt chi_adj chi_b''. */
static void acoustic_gradient(const R *hp, const R *chi, int irr, R xix_regular, const R *jac9,
                              const R *rhostore_e, R *gx, R *gy, R *gz)
{
    R t1[125], t2[125], t3[125];
    grad_ref(hp, chi, t1, t2, t3);
    for (int p = 0; p < 125; p++) {
        R rho_invl = (R)1 / rhostore_e[p];
        if (irr) {
            const R *J = jac9 + (int64_t)(irr - 1) * 9 * 125;
            R xixl = J[0 * 125 + p], xiyl = J[1 * 125 + p], xizl = J[2 * 125 + p];
            R etaxl = J[3 * 125 + p], etayl = J[4 * 125 + p], etazl = J[5 * 125 + p];
            R gaxl = J[6 * 125 + p], gayl = J[7 * 125 + p], gazl = J[8 * 125 + p];
            gx[p] = (t1[p] * xixl + t2[p] * etaxl + t3[p] * gaxl) * rho_invl;
            gy[p] = (t1[p] * xiyl + t2[p] * etayl + t3[p] * gayl) * rho_invl;
            gz[p] = (t1[p] * xizl + t2[p] * etazl + t3[p] * gazl) * rho_invl;
        } else {
            gx[p] = t1[p] * xix_regular * rho_invl;
            gy[p] = t2[p] * xix_regular * rho_invl;
            gz[p] = t3[p] * xix_regular * rho_invl;
        }
    }
}

/* receivers in acoustic elements: compute_seismograms.f90:127-144 with compute_interpolated_dva_acoust
 * (compute_interpolated_dva.f90:219-328), FP64 sums in (i, j, k) order:
 *   pressure != 0: out[r*3] = -sum h field(ibool) (field = potential with the pressure trick, else
 *                  potential_dot_dot), out[r*3 + 1..2] = 0;
 *   pressure == 0: (1/rho) grad(field) at the element's GLL points (compute_gradient_in_acoustic),
 *                  interpolated and rotated by nu[r*9 + 3*a + b]. Receivers in other domains are skipped. */
void FN(orc_seismo_acoustic)(int nrec, const int *rec_ispec, const double *hxir, const double *hetar,
                             const double *hgammar, const double *nu, const int *ibool,
                             const unsigned char *is_acoustic, const int *irregular, R xix_regular, const R *jac9,
                             const R *hprime, const R *rhostore, const R *field, int pressure, R *out)
{
    R chi[125], gx[125], gy[125], gz[125];
    for (int r = 0; r < nrec; r++) {
        int e = rec_ispec[r];
        if (!is_acoustic[e]) continue;
        const int *ib = ibool + (int64_t)e * 125;
        for (int p = 0; p < 125; p++) chi[p] = field[ib[p]];
        double dx = 0, dy = 0, dz = 0;
        if (pressure) {
            for (int k = 0; k < 5; k++)
                for (int j = 0; j < 5; j++)
                    for (int i = 0; i < 5; i++) {
                        double h = hxir[r * 5 + i] * hetar[r * 5 + j] * hgammar[r * 5 + k];
                        dx = dx - h * (double)chi[i + 5 * j + 25 * k];
                    }
            out[r * 3 + 0] = (R)dx;
            out[r * 3 + 1] = 0;
            out[r * 3 + 2] = 0;
            continue;
        }
        acoustic_gradient(hprime, chi, irregular[e], xix_regular, jac9, rhostore + (int64_t)e * 125, gx, gy, gz);
        for (int k = 0; k < 5; k++)
            for (int j = 0; j < 5; j++)
                for (int i = 0; i < 5; i++) {
                    int p = i + 5 * j + 25 * k;
                    double h = hxir[r * 5 + i] * hetar[r * 5 + j] * hgammar[r * 5 + k];
                    dx = dx + h * (double)gx[p];
                    dy = dy + h * (double)gy[p];
                    dz = dz + h * (double)gz[p];
                }
        const double *n = nu + (int64_t)r * 9;
        out[r * 3 + 0] = (R)(n[0] * dx + n[1] * dy + n[2] * dz);
        out[r * 3 + 1] = (R)(n[3] * dx + n[4] * dy + n[5] * dz);
        out[r * 3 + 2] = (R)(n[6] * dx + n[7] * dy + n[8] * dz);
    }
}

void FN(orc_compute_kernels_ac)(int nelem, const int *elems, const int *ibool, const int *irregular, R xix_regular,
                                const R *jac9, const R *hprime, const R *rhostore, const R *pot, const R *b_pot,
                                const R *b_potdd, R deltat, R *rho_ac_kl, R *kappa_ac_kl)
{
    R chi[125], bchi[125], ax[125], ay[125], az[125], bx[125], by[125], bz[125];
    for (int ie = 0; ie < nelem; ie++) {
        int e = elems[ie];
        const int *ib = ibool + (int64_t)e * 125;
        for (int p = 0; p < 125; p++) {
            bchi[p] = b_pot[ib[p]];
            chi[p] = pot[ib[p]];
        }
        const R *rho_e = rhostore + (int64_t)e * 125;
        acoustic_gradient(hprime, bchi, irregular[e], xix_regular, jac9, rho_e, bx, by, bz);
        acoustic_gradient(hprime, chi, irregular[e], xix_regular, jac9, rho_e, ax, ay, az);
        for (int p = 0; p < 125; p++) {
            int64_t ep = (int64_t)e * 125 + p;
            R rhol = rho_e[p];
            R a1 = rhol * ax[p], a2 = rhol * ay[p], a3 = rhol * az[p];
            R b1 = rhol * bx[p], b2 = rhol * by[p], b3 = rhol * bz[p];
            rho_ac_kl[ep] = rho_ac_kl[ep] + deltat * (a1 * b1 + a2 * b2 + a3 * b3);
            kappa_ac_kl[ep] = kappa_ac_kl[ep] + deltat * pot[ib[p]] * b_potdd[ib[p]];
        }
    }
}

/* max norm of a vector field and NaN flag: check_stability.f90 (max over points of |u|) */
R FN(orc_max_norm)(int64_t nglob, const R *f)
{
    R m = 0;
    for (int64_t i = 0; i < nglob; i++) {
        R v = (R)sqrt(f[3 * i] * f[3 * i] + f[3 * i + 1] * f[3 * i + 1] + f[3 * i + 2] * f[3 * i + 2]);
        if (v > m || v != v) m = v;
    }
    return m;
}

/* ------------------------------------------------------------------------- */
/* C-PML (elastic). Arrays of the C-PML elements in the reference's Fortran layouts:
 *   pml_displ[(ic*125 + p)*3 + c]             PML_displ_old/new(NDIM,NGLLX,NGLLY,NGLLZ,NSPEC_CPML)
 *   rm3[a][(s*ncpml + ic)*125 + p]            rmemory_du?_d?l_?(NGLLX,NGLLY,NGLLZ,NSPEC_CPML,3), a = 0..8:
 *       dux_dxl_x, duy_dxl_y, duz_dxl_z, dux_dyl_x, duy_dyl_y, duz_dyl_z, dux_dzl_x, duy_dzl_y, duz_dzl_z
 *   rm1[b][ic*125 + p]                        rmemory_*(NGLLX,NGLLY,NGLLZ,NSPEC_CPML), b = 0..11:
 *       duz_dzl_y, duz_dyl_y, duy_dzl_z, duy_dyl_z, duz_dzl_x, duz_dxl_x, dux_dzl_z, dux_dxl_z,
 *       duy_dyl_x, duy_dxl_x, dux_dyl_y, dux_dxl_y
 *   rdisp[((s*ncpml + ic)*125 + p)*3 + c]     rmemory_displ_elastic(NDIM,NGLLX,NGLLY,NGLLZ,NSPEC_CPML,3)
 *   coef_alpha/beta[(ic*125 + p)*9 + t], coef_abar[(..)*5 + t], coef_strain[(..)*18 + t]
 *       pml_convolution_coef_*(t,i,j,k,ispec_CPML)
 * Each array of rm3 / rm1 is passed as its own pointer (a pointer table). */

/* update_displ_elastic_PML (update_displacement_scheme.f90:417-455):
 * PML_displ = displ + deltatover2 (1 - 2 theta) veloc + deltatsqover2 (1 - theta) accel, theta = 1/8 */
void FN(orc_pml_store)(int ncpml, const int *cpml_to_spec, const int *ibool, const R *displ, const R *veloc,
                       const R *accel, R deltatover2, R deltatsqover2, R *pml_displ)
{
    const R theta = (R)0.125;
    const R c1 = deltatover2 * ((R)1 - (R)2 * theta);
    const R c2 = deltatsqover2 * ((R)1 - theta);
    for (int ic = 0; ic < ncpml; ic++) {
        const int e = cpml_to_spec[ic];
        for (int p = 0; p < 125; p++) {
            const int64_t g = ibool[(int64_t)e * 125 + p];
            for (int c = 0; c < 3; c++)
                pml_displ[((int64_t)ic * 125 + p) * 3 + c] =
                    displ[3 * g + c] + c1 * veloc[3 * g + c] + c2 * accel[3 * g + c];
        }
    }
}

/* pml_impose_boundary_condition_elastic (pml_compute_accel_contribution.f90:654-698): Dirichlet
 * condition on the outer faces of the C-PML layers (absorbing faces of elastic C-PML elements). */
void FN(orc_pml_dirichlet)(int nfaces, const int *face_ispec, const int *face_ijk, const int *ibool,
                           const unsigned char *is_elastic, const int *spec_to_cpml, R *displ, R *veloc, R *accel,
                           R *pml_old, R *pml_new)
{
    for (int f = 0; f < nfaces; f++) {
        const int e = face_ispec[f];
        if (!is_elastic[e] || !spec_to_cpml[e]) continue;
        const int ic = spec_to_cpml[e] - 1;
        for (int q = 0; q < 25; q++) {
            const int *ijk = face_ijk + ((int64_t)f * 25 + q) * 3;
            const int p = ijk[0] + 5 * ijk[1] + 25 * ijk[2];
            const int64_t g = ibool[(int64_t)e * 125 + p];
            for (int c = 0; c < 3; c++) {
                accel[3 * g + c] = 0;
                veloc[3 * g + c] = 0;
                displ[3 * g + c] = 0;
                pml_old[((int64_t)ic * 125 + p) * 3 + c] = 0;
                pml_new[((int64_t)ic * 125 + p) * 3 + c] = 0;
            }
        }
    }
}

/* r <- coef0 r + new coef1 + old coef2 (second-order recursive convolution, the expressions of
 * pml_compute_memory_variables.f90:152-296 and pml_compute_accel_contribution.f90:88-121) */
static inline R pml_conv(R r, R nw, R od, const R *c) { return c[0] * r + nw * c[1] + od * c[2]; }

/* C-PML element forces: compute_forces_viscoelastic_PML (compute_forces_viscoelastic.F90:969-1748)
 * with pml_compute_memory_variables_elastic (pml_compute_memory_variables.f90:29-415) and
 * pml_compute_accel_contribution_elastic (pml_compute_accel_contribution.f90:29-146). elems lists the
 * elastic elements in processing order; those with spec_to_cpml[e] == 0 are skipped. */
void FN(orc_forces_elastic_pml)(int nelem, const int *elems, const int *spec_to_cpml, int ncpml, const int *ibool,
                                const int *irregular, R xix_regular, R jacobian_regular, const R *jac9,
                                const R *jacobian, const R *hprime, const R *hprimewgll, const R *wgllwgll,
                                const R *wgll_cube, const R *kappastore, const R *mustore, const R *rhostore,
                                const R *displ, const R *veloc, R *accel, const R *pml_old, const R *pml_new,
                                R *const *rm3, R *const *rm1, R *rdisp, const R *coef_alpha, const R *coef_beta,
                                const R *coef_abar, const R *coef_strain)
{
    const R FOUR_THIRDS = (R)4 / (R)3;
    for (int ie = 0; ie < nelem; ie++) {
        const int e = elems[ie];
        if (!spec_to_cpml[e]) continue;
        const int ic = spec_to_cpml[e] - 1;
        const int *ib = ibool + (int64_t)e * 125;
        R f[9][125], t[27][125];  /* fields u, old, new (x,y,z each); their reference gradients */
        for (int p = 0; p < 125; p++) {
            const int64_t g = ib[p];
            for (int c = 0; c < 3; c++) {
                f[c][p] = displ[3 * g + c];
                f[3 + c][p] = pml_old[((int64_t)ic * 125 + p) * 3 + c];
                f[6 + c][p] = pml_new[((int64_t)ic * 125 + p) * 3 + c];
            }
        }
        for (int q = 0; q < 9; q++) grad_ref(hprime, f[q], t[3 * q], t[3 * q + 1], t[3 * q + 2]);
        const int irr = irregular[e];
        const R *J = irr ? jac9 + (int64_t)(irr - 1) * 9 * 125 : 0;
        R s1[9][125];  /* stress products tempx1, tempy1, tempz1, tempx2, ..., tempz3 */
        R cpml[3][125];
        for (int p = 0; p < 125; p++) {
            /* derivatives D[field][comp][dir]: d u_comp / d x_dir of u (0), old (1), new (2) */
            R D[3][3][3];
            R xixl = 0, xiyl = 0, xizl = 0, etaxl = 0, etayl = 0, etazl = 0, gaxl = 0, gayl = 0, gazl = 0;
            if (irr) {
                xixl = J[0 * 125 + p]; xiyl = J[1 * 125 + p]; xizl = J[2 * 125 + p];
                etaxl = J[3 * 125 + p]; etayl = J[4 * 125 + p]; etazl = J[5 * 125 + p];
                gaxl = J[6 * 125 + p]; gayl = J[7 * 125 + p]; gazl = J[8 * 125 + p];
            }
            for (int fl = 0; fl < 3; fl++)
                for (int c = 0; c < 3; c++) {
                    const R a1 = t[3 * (3 * fl + c)][p], a2 = t[3 * (3 * fl + c) + 1][p], a3 = t[3 * (3 * fl + c) + 2][p];
                    if (irr) {
                        D[fl][c][0] = xixl * a1 + etaxl * a2 + gaxl * a3;
                        D[fl][c][1] = xiyl * a1 + etayl * a2 + gayl * a3;
                        D[fl][c][2] = xizl * a1 + etazl * a2 + gazl * a3;
                    } else {
                        D[fl][c][0] = xix_regular * a1;
                        D[fl][c][1] = xix_regular * a2;
                        D[fl][c][2] = xix_regular * a3;
                    }
                }
            const int64_t lp = (int64_t)ic * 125 + p;
            const R *ca = coef_alpha + lp * 9, *cb = coef_beta + lp * 9, *A = coef_strain + lp * 18;
            const R *cax = ca, *cay = ca + 3, *caz = ca + 6;
            const R *cbx = cb, *cby = cb + 3, *cbz = cb + 6;
            /* the three-coefficient families: derivative direction d, coefficient sets per sls index */
            const R *fam[3][3] = {{caz, cay, cbx}, {cax, caz, cby}, {cax, cay, cbz}};
            R r3[9][3];
            for (int d = 0; d < 3; d++)
                for (int c = 0; c < 3; c++) {
                    R *arr = rm3[3 * d + c];
                    for (int s = 0; s < 3; s++) {
                        R *r = arr + ((int64_t)s * ncpml + ic) * 125 + p;
                        *r = pml_conv(*r, D[2][c][d], D[1][c][d], fam[d][s]);
                        r3[3 * d + c][s] = *r;
                    }
                }
            /* single-coefficient memory variables: (array, comp, dir, coefficient set) */
            static const int one[12][2] = {{2, 2}, {2, 1}, {1, 2}, {1, 1}, {2, 2}, {2, 0},
                                           {0, 2}, {0, 0}, {1, 1}, {1, 0}, {0, 1}, {0, 0}};
            R r1[12];
            for (int b = 0; b < 12; b++) {
                const R *cs = b < 4 ? cax : (b < 8 ? cay : caz);
                R *r = rm1[b] + lp;
                *r = pml_conv(*r, D[2][one[b][0]][one[b][1]], D[1][one[b][0]][one[b][1]], cs);
                r1[b] = *r;
            }
            /* A6..A23 = A[0..17] */
            const R (*C)[3] = D[0];
            const R duxdxl_x = A[0] * C[0][0] + A[1] * r3[0][0] + A[2] * r3[0][1] + A[3] * r3[0][2];
            const R duxdyl_x = A[4] * C[0][1] + A[5] * r3[3][0] + A[6] * r3[3][1] + A[7] * r3[3][2];
            const R duxdzl_x = A[8] * C[0][2] + A[9] * r3[6][0] + A[10] * r3[6][1] + A[11] * r3[6][2];
            const R duzdzl_x = A[14] * C[2][2] + A[15] * r1[4];
            const R duzdxl_x = A[14] * C[2][0] + A[15] * r1[5];
            const R duydyl_x = A[16] * C[1][1] + A[17] * r1[8];
            const R duydxl_x = A[16] * C[1][0] + A[17] * r1[9];
            const R duydxl_y = A[0] * C[1][0] + A[1] * r3[1][0] + A[2] * r3[1][1] + A[3] * r3[1][2];
            const R duydyl_y = A[4] * C[1][1] + A[5] * r3[4][0] + A[6] * r3[4][1] + A[7] * r3[4][2];
            const R duydzl_y = A[8] * C[1][2] + A[9] * r3[7][0] + A[10] * r3[7][1] + A[11] * r3[7][2];
            const R duzdzl_y = A[12] * C[2][2] + A[13] * r1[0];
            const R duzdyl_y = A[12] * C[2][1] + A[13] * r1[1];
            const R duxdyl_y = A[16] * C[0][1] + A[17] * r1[10];
            const R duxdxl_y = A[16] * C[0][0] + A[17] * r1[11];
            const R duzdxl_z = A[0] * C[2][0] + A[1] * r3[2][0] + A[2] * r3[2][1] + A[3] * r3[2][2];
            const R duzdyl_z = A[4] * C[2][1] + A[5] * r3[5][0] + A[6] * r3[5][1] + A[7] * r3[5][2];
            const R duzdzl_z = A[8] * C[2][2] + A[9] * r3[8][0] + A[10] * r3[8][1] + A[11] * r3[8][2];
            const R duydzl_z = A[12] * C[1][2] + A[13] * r1[2];
            const R duydyl_z = A[12] * C[1][1] + A[13] * r1[3];
            const R duxdzl_z = A[14] * C[0][2] + A[15] * r1[6];
            const R duxdxl_z = A[14] * C[0][0] + A[15] * r1[7];
            const R kappal = kappastore[(int64_t)e * 125 + p], mul = mustore[(int64_t)e * 125 + p];
            const R l2m = kappal + FOUR_THIRDS * mul;
            const R lam = l2m - (R)2 * mul;
            const R sxx = l2m * duxdxl_x + lam * duydyl_x + lam * duzdzl_x;
            const R syx = mul * duxdyl_x + mul * duydxl_x;
            const R szx = mul * duzdxl_x + mul * duxdzl_x;
            const R sxy = mul * duxdyl_y + mul * duydxl_y;
            const R syy = lam * duxdxl_y + l2m * duydyl_y + lam * duzdzl_y;
            const R szy = mul * duzdyl_y + mul * duydzl_y;
            const R sxz = mul * duzdxl_z + mul * duxdzl_z;
            const R syz = mul * duzdyl_z + mul * duydzl_z;
            const R szz = lam * duxdxl_z + lam * duydyl_z + l2m * duzdzl_z;
            R jl;
            if (irr) {
                jl = jacobian[(int64_t)(irr - 1) * 125 + p];
                s1[0][p] = jl * (sxx * xixl + syx * xiyl + szx * xizl);
                s1[1][p] = jl * (sxy * xixl + syy * xiyl + szy * xizl);
                s1[2][p] = jl * (sxz * xixl + syz * xiyl + szz * xizl);
                s1[3][p] = jl * (sxx * etaxl + syx * etayl + szx * etazl);
                s1[4][p] = jl * (sxy * etaxl + syy * etayl + szy * etazl);
                s1[5][p] = jl * (sxz * etaxl + syz * etayl + szz * etazl);
                s1[6][p] = jl * (sxx * gaxl + syx * gayl + szx * gazl);
                s1[7][p] = jl * (sxy * gaxl + syy * gayl + szy * gazl);
                s1[8][p] = jl * (sxz * gaxl + syz * gayl + szz * gazl);
            } else {
                jl = jacobian_regular;
                s1[0][p] = jl * sxx * xix_regular;
                s1[1][p] = jl * sxy * xix_regular;
                s1[2][p] = jl * sxz * xix_regular;
                s1[3][p] = jl * syx * xix_regular;
                s1[4][p] = jl * syy * xix_regular;
                s1[5][p] = jl * syz * xix_regular;
                s1[6][p] = jl * szx * xix_regular;
                s1[7][p] = jl * szy * xix_regular;
                s1[8][p] = jl * szz * xix_regular;
            }
            /* pml_compute_accel_contribution_elastic */
            const R *ab = coef_abar + lp * 5;
            const R *cs3[3] = {cax, cay, caz};
            const int64_t g = ib[p];
            const R w = wgll_cube[p] * rhostore[(int64_t)e * 125 + p] * jl;
            for (int c = 0; c < 3; c++) {
                R rr[3];
                for (int s = 0; s < 3; s++) {
                    R *r = rdisp + (((int64_t)s * ncpml + ic) * 125 + p) * 3 + c;
                    *r = pml_conv(*r, f[6 + c][p], f[3 + c][p], cs3[s]);
                    rr[s] = *r;
                }
                cpml[c][p] = w * (ab[0] * veloc[3 * g + c] + ab[1] * displ[3 * g + c] + ab[2] * rr[0] + ab[3] * rr[1] +
                                  ab[4] * rr[2]);
            }
        }
        R n1[125], n2[125], n3[125];
        for (int c = 0; c < 3; c++) {
            div_ref(hprimewgll, s1[c], s1[3 + c], s1[6 + c], n1, n2, n3);
            for (int k = 0; k < 5; k++)
                for (int j = 0; j < 5; j++)
                    for (int i = 0; i < 5; i++) {
                        const int p = i + 5 * j + 25 * k;
                        const R fac1 = wgllwgll[j * 5 + k], fac2 = wgllwgll[i * 5 + k], fac3 = wgllwgll[i * 5 + j];
                        const int64_t g = ib[p];
                        accel[3 * g + c] = accel[3 * g + c] - (fac1 * n1[p] + fac2 * n2[p] + fac3 * n3[p] + cpml[c][p]);
                    }
        }
    }
}
