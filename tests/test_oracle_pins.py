"""Hand-derived rows that pin the oracle's advection details which closed
forms (CFL-1 shifts, conservation, eigenmodes) leave open.  Each case is a
row of 3-4 cells with dyadic inputs whose result was worked out by hand
(the working is in the comments), chosen so that the plausible alternative
reading of the reference gives a different number:

* fss at interface I uses dt/kappa of cell I (src/AdvectionHorz_DCU.f90:
  155-156, :230), not of cell I-1 or an average of the two;
* the limiter is skipped for |dq| <= 1e-10, including |dq| == 1e-10, and
  applied just above it (:183);
* theta is clamped to sign(3, dqu*dq) when |dqu| > 3|dq|, with the sign of
  the product (:200-205), shown with the Fromm limiter (which, unlike
  superbee, separates theta = 3 from the unclamped value and -3 from +3);
* the topographic terms 0.5*vel_I*(sig_I+sig_{I+1})*q_I*D_I on the right
  fluctuation and 0.5*vel_{I-1}*(sig_{I-1}+sig_I)*q_{I-1}*D_{I-1} on the
  left one (:250-257), on a row with non-uniform vel, sigma and D;
* advect_z with kmin > 1 adds update_cc(ncells+1), the update of an
  interior cell, to the top outflow (src/AdvectionVert_DCU.f90:309-311),
  records no deposit, and zeroes the cells below kmin (the zero-filled ts1
  copied back, :151, :318-319).

The same rows run through the HIP library in tests/test_gpu_pins.py."""
import numpy as np
import pytest

from volcano_rows import advect_x_row, advect_z_column

EPS = 1.0e-10


def check(res, q_expect, of_lo, of_hi):
    q, lo, hi = res
    assert np.array_equal(q, np.array(q_expect)), (q, q_expect)
    assert lo == of_lo, (lo, of_lo)
    assert hi == of_hi, (hi, of_hi)


# -- dt/kappa of cell I inside fss -----------------------------------------
# Row of 4 cells, u = vel*sigma = 1 everywhere, dt = 1, superbee.
# q = [4, 8, 8, 2] (ghosts 0): dq_1..dq_5 = 4, 4, 0, -6, -2.
# kappa (cells 1..5) = 4, 2, 8, 4, 8 (cell 5 is the ghost past the row).
# fss_I = 0.5*u*(1 - dt/kappa_I*u)*ldq_I:
#   I=1: theta = dq_0/dq_1 = 0 -> phi = 0 -> 0
#   I=2: theta = 1 -> phi = 1, ldq = 4 -> 2*(1 - 1/2) = 1
#   I=3: |dq| = 0 -> skipped -> 0;  I=4: theta = 0 -> 0
#   I=5: theta = -6/-2 = 3 -> phi = 2, ldq = -4 -> -2*(1 - 1/8) = -7/4
# update_c = -dt/kappa_c*(u*dq_c + fss_{c+1} - fss_c):
#   c=1: -(4 + 1)/4 = -5/4  -> 11/4
#   c=2: -(4 - 1)/2 = -3/2  -> 13/2
#   c=3: 0                  -> 8
#   c=4: -(-6 - 7/4)/4 = 31/16 -> 63/16
#   ghost 5: -(fsr_5 + 0 - fss_5)/8 = -(-2 + 7/4)/8 = 1/32 (east outflow)
# With dt/kappa_{I-1}: fss_2 = 3/2, fss_5 = -3/2 and q1 = 21/8 instead.
def test_fss_uses_kappa_of_cell_I(oracle):
    res = advect_x_row(oracle, q=[4, 8, 8, 2], vel=[1, 1, 1, 1], sig=[1] * 8,
                       kap=[4, 4, 4, 2, 8, 4, 8, 8], dt=1.0, limiter=5)
    check(res, [11 / 4, 13 / 2, 8, 63 / 16], 0.0, 1 / 32)


# -- the |dq| <= EPS_THRESH skip ------------------------------------------
# Lax-Wendroff (ldq = dq), 2 cells, u = 1, kappa = 2, dt = 1: dt/kappa = 1/2,
# fss_I = 0.5*(1 - 1/2)*dq_I = dq_I/4 unless skipped.  q = [e, 0].
#  e = 1e-10 (== EPS_THRESH): every |dq| <= EPS -> skipped ->
#     q1 = e - e/2 = e/2, q2 = 0 + e/2 = e/2 (a strict '<' would give 3e/4)
#  e = 2^-33 (just above): fss_1 = e/4, fss_2 = -e/4, fss_3 = 0 ->
#     q1 = e - (e - e/4 - e/4)/2 = 3e/4, q2 = -(-e + e/4)/2 = 3e/8
def test_eps_skip_at_and_above_threshold(oracle):
    e = EPS
    res = advect_x_row(oracle, q=[e, 0.0], vel=[1, 1], sig=[1] * 6, kap=[2] * 6, dt=1.0, limiter=1)
    check(res, [e / 2, e / 2], 0.0, 0.0)
    e = 2.0 ** -33
    assert e > EPS
    res = advect_x_row(oracle, q=[e, 0.0], vel=[1, 1], sig=[1] * 6, kap=[2] * 6, dt=1.0, limiter=1)
    check(res, [3 * e / 4, 3 * e / 8], -e / 8, 0.0)


# -- the theta clamp, both signs --------------------------------------------
# Fromm (phi = (1+theta)/2), 3 cells, u = 1, kappa = 4, dt = 1:
# fss_I = 0.5*(1 - 1/4)*ldq = 3/8*ldq.
# Row A, q = [4, 5, 1]: dq_1..dq_4 = 4, 1, -4, -1.
#   I=1 theta 0 -> ldq 2 -> 3/4;  I=2 |4| > 3*1 -> theta +3 -> ldq 2 -> 3/4
#   (unclamped theta 4 would give 15/16);  I=3 theta -1/4 -> ldq -3/2 -> -9/16
#   I=4 |-4| > 3 -> theta sign(3, 4) = +3 -> ldq -2 -> -3/4
#   c=1: -(4 + 3/4 - 3/4)/4 = -1 -> 3
#   c=2: -(1 - 9/16 - 3/4)/4 = 5/64 -> 325/64
#   c=3: -(-4 - 3/4 + 9/16)/4 = 67/64 -> 131/64
#   ghost 4: -(-1 - (-3/4))/4 = 1/16; ghost 0: -(fss_1)/4 = -3/16
# Row B, q = [4, 0, 1]: dq = 4, -4, 1, -1.
#   I=1 -> 3/4; I=2 theta -1 -> phi 0 -> 0
#   I=3 |-4| > 3 -> theta sign(3, -4) = -3 -> phi -1 -> ldq -1 -> -3/8
#   (theta +3 would give +3/4);  I=4 theta -1 -> 0
#   c=1: -(4 + 0 - 3/4)/4 = -13/16 -> 51/16
#   c=2: -(-4 - 3/8 - 0)/4 = 35/32 -> 35/32
#   c=3: -(1 + 0 + 3/8)/4 = -11/32 -> 21/32
#   ghost 4: -(-1 - 0)/4 = 1/4; ghost 0: -3/16
def test_theta_clamp_both_signs(oracle):
    res = advect_x_row(oracle, q=[4, 5, 1], vel=[1, 1, 1], sig=[1] * 7, kap=[4] * 7, dt=1.0, limiter=3)
    check(res, [3.0, 325 / 64, 131 / 64], -3 / 16, 1 / 16)
    res = advect_x_row(oracle, q=[4, 0, 1], vel=[1, 1, 1], sig=[1] * 7, kap=[4] * 7, dt=1.0, limiter=3)
    check(res, [51 / 16, 35 / 32, 21 / 32], -3 / 16, 1 / 4)


# -- the topographic DelDonD terms ------------------------------------------
# No limiter (fss = 0), 3 cells, kappa = 4, dt = 1 (dt/kappa = 1/4).
# vel (cells -1..5, ghosts = edge values) = 1, 1, 1, 2, 1, 1, 1
# sigma (cells -1..5)                    = 1, 1, 2, 1, 2, 1, 1
# D = DelDxonD (cells 0..4; 0 beyond)    = 1/2, 1/4, -1/2, 1/8, 3/4
# q = [4, 8, 2].  usig_I = (vel_{I-1}+vel_I)/2*sig_I, I = 0..4:
#   1, 2, 3/2, 3, 1 (0 outside), all >= 0, so divu_m = 0 and
#   divu_p(I) = u_I - u_{I-1}.  tf_I = vel_I*(sig_I + sig_{I+1})/2, I = 0..4:
#   3/2, 3/2, 3, 3/2, 1.
# fsr_I = u_I*dq_I + tf_I*q_I*D_I;  fsl_I = q_{I-1}*divu_p(I) - tf_{I-1}*q_{I-1}*D_{I-1}
#   fsr_1 = 8 + 3/2 = 19/2     fsl_2 = -2 - 3/2 = -7/2
#   fsr_2 = 6 - 12 = -6        fsl_3 = 12 + 12 = 24
#   fsr_3 = -18 + 3/8 = -141/8 fsl_4 = -4 - 3/8 = -35/8
#   fsr_4 = -2 (q_4 = 0)       fsl_1 = fsr_0 = 0 (q_0 = 0)
# update_c = -(fsl_{c+1} + fsr_c)/4: -3/2, -9/2, 11/2 -> q = 5/2, 7/2, 15/2;
# ghost 4: -(0 + fsr_4)/4 = 1/2 (east outflow); ghost 0: 0.
# With vel_I in place of vel_{I-1} in the left term, fsl_2 = -5 and q1 = 23/8.
def test_topography_terms(oracle):
    res = advect_x_row(oracle, q=[4, 8, 2], vel=[1, 2, 1], sig=[1, 1, 2, 1, 2, 1, 1],
                       kap=[4] * 7, dt=1.0, limiter=0, deld=[0, 0.5, 0.25, -0.5, 0.125, 0.75, 0])
    check(res, [5 / 2, 7 / 2, 15 / 2], 0.0, 0.5)


# -- advect_z top outflow with kmin > 1 ---------------------------------------
# One column of 5 cells, active kmin..kmax = 3..5 (ncells = 3), no limiter,
# vz = vh = 0, vf = -1 (settling), sigma_z = 1, kappa = 4, dt = 1.
# q(1..5) = 9, 6, 4, 8, 2; z ghosts 0.  usig_I = -1 for I = 2..6, 0 else.
# dq_I (I = 2..7) = -3, -2, 4, -6, -2, 0.  With u < 0: fsl_I = -dq_I,
# fsr_I = q_I*divu_m(I) = 0 (divu_m(6) = 0 - (-1) = 1 but q_6 = 0).
# update_c = -(fsl_{c+1} + fsr_c)/4, c = 2..6: -1/2, 1, -3/2, -1/2, 0.
# Interior 3..5: 4+1, 8-3/2, 2-1/2 = 5, 13/2, 3/2; cells 1, 2 (below kmin)
# are left zero; top outflow += update_cc(ncells+1) = update of cell 4 =
# -3/2 (not the ghost nz+1, whose update is 0); no deposit (kmin != 1).
def test_top_outflow_index_with_kmin_above_one(oracle):
    q, dep, top = advect_z_column(oracle, q=[9, 6, 4, 8, 2], vf=-1.0, sig=1.0, kap=4.0, dt=1.0,
                                  kmin=3, kmax=5, limiter=0)
    assert np.array_equal(q, np.array([0.0, 0.0, 5.0, 6.5, 1.5])), q
    assert dep == 0.0
    assert top == -1.5
