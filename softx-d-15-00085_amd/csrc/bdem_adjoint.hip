/*
 * bdem_adjoint.hip - reverse-mode (adjoint) kernels of the sphere step under
 * the reference pair law, for calibrating material parameters.
 *
 * The forward step is the library's own (k_step: contact scan, pair law
 * Collision_Response.cuh:374-494 with gravity, drum walls
 * Collision_Response.cuh:116-233 and 504-526, symplectic Euler
 * Integration.cuh:6-184).  A recorded step keeps its sorted input state, its
 * contact lists (sorted positions, reference order) and the permutation of
 * its sort; the adjoint of that step, given the cotangent of its output
 * state, produces the cotangent of its input state in the previous step's
 * order and the step's contribution to the parameter gradient.  The contact
 * set is frozen within a step (the force is not differentiable where a
 * contact opens or closes).
 *
 * Every vector-Jacobian product below is derived by hand from the forward
 * expressions: each forward line y = f(x) becomes the accumulation
 * gx += (df/dx)^T gy, visited in reverse order.  For a cross product
 * c = a x b: ga += b x gc and gb += gc x a.  For a unit vector n = u / |u|:
 * gu += (gn - (gn . n) n) / |u|.  For a projection u = V - (V . n) n:
 * gV += gu - (gu . n) n and gn -= (V . n) gu + (gu . n) V.
 */
#include "bdem_kernels.hpp"
#include "bdem_adjoint_k.hpp"

namespace bdem {

/* Reverse of the projection u = V - (V . n) n. */
__device__ __forceinline__ void proj_vjp(f3 V, f3 n, f3 gu, f3& gV, f3& gn) {
  gV = gV + (gu - dot(gu, n) * n);
  gn = gn - (dot(V, n) * gu + dot(gu, n) * V);
}

/* Reverse of n = u / |u| (|u| = len). */
__device__ __forceinline__ f3 unit_vjp(f3 n, float len, f3 gn) { return (gn - dot(gn, n) * n) * (1.0f / len); }

/* Reverse of pair_ref (the force and torque on A from B).  gF, gT: the
 * cotangents of that force and torque.  Accumulates the cotangents of both
 * bodies' position, velocity and spin and of the pair's Kn, Cn and kinetic
 * friction. */
__device__ __forceinline__ void pair_ref_vjp(const DevParams& P, f3 pa, float ra, f3 va, float ma, f3 wa, int ta,
                                             f3 pb, float rb, f3 vb, float mb, f3 wb, int tb, f3 gF, f3 gT,
                                             f3& gpa, f3& gva, f3& gwa, f3& gpb, f3& gvb, f3& gwb, float& gKn,
                                             float& gCn, float& gFk) {
  /* forward, as pair_ref and pair_resp compute it */
  const f3 rel = pa - pb;
  const float d = length(rel);
  const f3 N = rel / d;
  const float S = ra + rb;
  const float dis = fabsf(d - S) + 0.000010f;
  const f3 RA = (-1.0f * N) * ra;
  const f3 RB = N * rb;
  const f3 RVel = va - vb;
  const float Kn = P.pp_kn[ta][tb], Cn = P.pp_cn[ta][tb], Fk = P.pp_fk[ta][tb];
  const float vn = dot(RVel, N);
  const float sn = Kn * dis - Cn * vn;
  const f3 V_CPA = va + cross(wa, RA);
  const f3 V_CPB = vb + cross(wb, RB);
  const f3 Vt = (V_CPA - dot(V_CPA, N) * N) - (V_CPB - dot(V_CPB, N) * N);
  const float mag2 = dot(Vt, Vt);
  const float c = ((0.50f * (ma + mb)) / P.dt) * 0.050f;
  float velf = dot(RVel, RVel);
  const bool vpos = velf > 0.0f;
  velf = vpos ? __builtin_sqrtf(velf) : 1.0f;
  const float kappa = (mag2 > 0.0f) ? (c * Fk) / velf : 0.0f;   /* Tang = -kappa Vt */
  const f3 F = sn * N - kappa * Vt;
  /* reverse */
  f3 gRA = mk3(0.f, 0.f, 0.f), gRB = gRA, gN = gRA, gRVel = gRA;
  f3 gFt = gF;
  if (P.rotation) {                                    /* T = RA x F */
    gRA = gRA + cross(F, gT);
    gFt = gFt + cross(gT, RA);
  }
  const float gsn = dot(gFt, N);                       /* F = sn N - kappa Vt */
  gN = gN + sn * gFt;
  f3 gVt = mk3(0.f, 0.f, 0.f);
  if (mag2 > 0.0f) {
    gVt = (-kappa) * gFt;
    const float gkappa = -dot(Vt, gFt);
    gFk += gkappa * c / velf;
    if (vpos) {
      const float gvelf = -gkappa * kappa / velf;
      gRVel = gRVel + (gvelf / velf) * RVel;
    }
  }
  gKn += gsn * dis;                                    /* sn = Kn dis - Cn vn */
  gCn += -gsn * vn;
  const float gdis = gsn * Kn;
  const float gvn = -gsn * Cn;
  gRVel = gRVel + gvn * N;                             /* vn = RVel . N */
  gN = gN + gvn * RVel;
  /* Vt = P_N(V_CPA) - P_N(V_CPB) */
  f3 gVA = mk3(0.f, 0.f, 0.f), gVB = gVA;
  proj_vjp(V_CPA, N, gVt, gVA, gN);
  proj_vjp(V_CPB, N, (-1.0f) * gVt, gVB, gN);
  gva = gva + gVA;                                     /* V_CPA = va + wa x RA */
  gwa = gwa + cross(RA, gVA);
  gRA = gRA + cross(gVA, wa);
  gvb = gvb + gVB;                                     /* V_CPB = vb + wb x RB */
  gwb = gwb + cross(RB, gVB);
  gRB = gRB + cross(gVB, wb);
  gva = gva + gRVel;                                   /* RVel = va - vb */
  gvb = gvb - gRVel;
  gN = gN + (-ra) * gRA + rb * gRB;                    /* RA = -ra N, RB = rb N */
  const float gd = (d - S >= 0.0f ? 1.0f : -1.0f) * gdis;   /* dis = |d - S| + 1e-5 */
  const f3 grel = unit_vjp(N, d, gN) + gd * N;         /* N = rel / d, d = |rel| */
  gpa = gpa + grel;
  gpb = gpb - grel;
}

/* Reverse of sphere_cylinder + resp_moving (the rotating drum's wall). */
__device__ __forceinline__ void drum_wall_vjp(const DevParams& P, int t, f3 p, float r, f3 v, f3 w, float m,
                                              const bdem_Macro_Cylinder& cyl, f3 gF, f3 gT, f3& gp, f3& gv,
                                              f3& gw) {
  const f3 A = cvt(cyl.AXIS);
  const f3 Pvec = p * A - cvt(cyl.center_bot_cap);
  const float L = length(Pvec);
  const f3 u = -1.0f * Pvec;
  const float lu = length(u);
  const f3 n = u * (1.0f / lu);
  const f3 cp = p - n * r;
  const float Rd = P.wobj[0].cyl_geo.radius;
  const f3 q = cp - mk3(Rd, Rd, 0.0f);
  const float Om = P.mill_rot_vel;
  const f3 Vs = cross(mk3(0.0f, 0.0f, -Om), q);
  const f3 Rc = cp - p;
  const f3 RVel = v - Vs;
  const float vn = dot(RVel, n);
  const float dis = fabsf((L + r) - cyl.radius);
  const float Fnormal = P.s_kn[t] * dis;
  const float sn_full = Fnormal - P.s_cn[t] * vn;
  /* the forward's attractive test: the direction of Fnormal n - Cn (RVel . n) n against n */
  const f3 Ftr0 = Fnormal * n - P.s_cn[t] * (vn * n);
  const f3 Fdir = Ftr0 / length(Ftr0);
  const bool attractive = dot(Fdir, n) < 0.0f;
  const float sn = attractive ? Fnormal : sn_full;
  const f3 V_CPA = v + cross(w, Rc);
  const f3 Vt = (V_CPA - dot(V_CPA, n) * n) - (Vs - dot(Vs, n) * n);
  const float mag2 = dot(Vt, Vt);
  float velf = dot(RVel, RVel);
  const bool vpos = velf > 0.0f;
  velf = vpos ? __builtin_sqrtf(velf) : 1.0f;
  const float c = P.s_fk[t] * (m / P.dt) * 0.050f;
  const float kappa = mag2 > 0.0f ? c / velf : 0.0f;
  const f3 F = sn * n - kappa * Vt;
  /* reverse */
  f3 gRc = mk3(0.f, 0.f, 0.f), gn = gRc, gRVel = gRc, gVs = gRc;
  f3 gFt = gF;
  if (P.rotation) {
    gRc = gRc + cross(F, gT);
    gFt = gFt + cross(gT, Rc);
  }
  const float gsn = dot(gFt, n);
  gn = gn + sn * gFt;
  f3 gVt = mk3(0.f, 0.f, 0.f);
  if (mag2 > 0.0f) {
    gVt = (-kappa) * gFt;
    const float gkappa = -dot(Vt, gFt);
    if (vpos) gRVel = gRVel + ((-gkappa * kappa / velf) / velf) * RVel;
  }
  const float gdis = gsn * P.s_kn[t];
  if (!attractive) {
    const float gvn = -gsn * P.s_cn[t];
    gRVel = gRVel + gvn * n;
    gn = gn + gvn * RVel;
  }
  f3 gVA = mk3(0.f, 0.f, 0.f);
  proj_vjp(V_CPA, n, gVt, gVA, gn);
  proj_vjp(Vs, n, (-1.0f) * gVt, gVs, gn);
  gv = gv + gVA;                                       /* V_CPA = v + w x Rc */
  gw = gw + cross(Rc, gVA);
  gRc = gRc + cross(gVA, w);
  gv = gv + gRVel;                                     /* RVel = v - Vs */
  gVs = gVs - gRVel;
  f3 gcp = cross(gVs, mk3(0.0f, 0.0f, -Om));           /* Vs = a x q: gq = gVs x a */
  gcp = gcp + gRc;                                     /* Rc = cp - p */
  gp = gp - gRc;
  gp = gp + gcp;                                       /* cp = p - r n */
  gn = gn + (-r) * gcp;
  const float gL = ((L + r) - cyl.radius >= 0.0f ? 1.0f : -1.0f) * gdis;
  const f3 gu = unit_vjp(n, lu, gn);
  f3 gP = (-1.0f) * gu + (L > 0.0f ? (gL / L) * Pvec : mk3(0.f, 0.f, 0.f));
  gp = gp + gP * A;                                    /* Pvec = p * AXIS - c_bot */
}

/* Reverse of sphere_plane (infinite plane) + resp_endplate. */
__device__ __forceinline__ void endplate_vjp(const DevParams& P, int t, f3 nrm, f3 gF, f3& gp, f3& gv) {
  const float gs = dot(gF, nrm);             /* F = (Kn dis - Cn (v . n)) n */
  gp = gp - (P.s_kn[t] * gs) * nrm;          /* dis = -(n . (p - r n - c)) */
  gv = gv - (P.s_cn[t] * gs) * nrm;
}

/* Reverse of wall_forces for the walls the adjoint supports: cylinder world
 * objects with the moving drum surface and flat end plates.  The forward
 * assigns the force at each cylinder or cap hit (the last one stands) and
 * adds the cylinder's torque. */
__device__ __forceinline__ void wall_vjp(const DevParams& P, int t, f3 p, float r, f3 v, f3 w, float m, f3 gF, f3 gT,
                                         f3& gp, f3& gv, f3& gw) {
  int last = -1;        /* which contribution's force survives: 0 cylinder, 1 top, 2 bottom */
  int last_k = -1;
  for (int k = 0; k < P.num_wobj; ++k) {
    const bdem_Macro_Cylinder& cyl = P.wobj[k].cyl_geo;
    if (sphere_cylinder(p, r, cyl).hit) { last = 0; last_k = k; }
    if (cyl.has_top_cap && sphere_plane(p, r, cvt(cyl.normal_top_cap), cvt(cyl.center_top_cap), nullptr, nullptr).hit) {
      last = 1; last_k = k;
    }
    if (cyl.has_bot_cap && sphere_plane(p, r, cvt(cyl.normal_bot_cap), cvt(cyl.center_bot_cap), nullptr, nullptr).hit) {
      last = 2; last_k = k;
    }
  }
  for (int k = 0; k < P.num_wobj; ++k) {
    const bdem_Macro_Cylinder& cyl = P.wobj[k].cyl_geo;
    if (sphere_cylinder(p, r, cyl).hit) {
      const f3 gFk = (last == 0 && last_k == k) ? gF : mk3(0.f, 0.f, 0.f);
      drum_wall_vjp(P, t, p, r, v, w, m, cyl, gFk, gT, gp, gv, gw);
    }
  }
  if (last == 1) endplate_vjp(P, t, cvt(P.wobj[last_k].cyl_geo.normal_top_cap), gF, gp, gv);
  if (last == 2) endplate_vjp(P, t, cvt(P.wobj[last_k].cyl_geo.normal_bot_cap), gF, gp, gv);
}

/* ------------------------------------------------------------------------ */
/* Observables: per-block moments of the state (mass, first and second     */
/* moments of x and y), in double, one row of blocks per step.             */
/* ------------------------------------------------------------------------ */
__global__ __launch_bounds__(ADJ_THREADS) void k_adj_moments(int n, const float4* __restrict__ pos,
                                                             const float4* __restrict__ vel,
                                                             const float4* __restrict__ omg,
                                                             double* __restrict__ out) {
  __shared__ double red[ADJ_MOMENTS][ADJ_THREADS / 64];
  const int i = blockIdx.x * ADJ_THREADS + threadIdx.x;
  double m[ADJ_MOMENTS] = {0, 0, 0, 0, 0, 0};
  if (i < n && packed_pid(omg[i].w) >= 0) {
    const double x = pos[i].x, y = pos[i].y, mm = vel[i].w;
    m[0] = mm; m[1] = mm * x; m[2] = mm * y; m[3] = mm * x * x; m[4] = mm * y * y; m[5] = mm * x * y;
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < ADJ_MOMENTS; ++k) {
    double s = m[k];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
    if (lane == 0) red[k][wave] = s;
  }
  __syncthreads();
  if (threadIdx.x < ADJ_MOMENTS) {
    double s = 0.0;
    for (int w = 0; w < ADJ_THREADS / 64; ++w) s += red[threadIdx.x][w];
    out[(size_t)blockIdx.x * ADJ_MOMENTS + threadIdx.x] = s;
  }
}

/* ------------------------------------------------------------------------ */
/* Reverse step, part 1 (per sorted position): recompute the forces and the */
/* integration, add the loss's cotangent of the output position, reverse    */
/* the integrator.  lam_out: cotangent of the step's output (x, v, w) in    */
/* sorted order; writes the cotangents of the total force and torque and    */
/* the integrator's direct part of the input cotangent.                     */
/* ------------------------------------------------------------------------ */
__global__ __launch_bounds__(ADJ_THREADS) void k_adj_back1(const DevParams* __restrict__ Pp, AdjStep st,
                                                           const float4* __restrict__ lam_out,
                                                           float4* __restrict__ lamF, float4* __restrict__ lamIn,
                                                           LossCoef g, double* __restrict__ part_rr) {
  const DevParams& P = *Pp;
  __shared__ double red[ADJ_THREADS / 64];
  const int i = blockIdx.x * ADJ_THREADS + threadIdx.x;
  double grr = 0.0;
  if (i < st.n) {
    const float4 pr = st.pos[i], vm = st.vel[i], ot = st.omg[i];
    const int pid = packed_pid(ot.w), ta = packed_type(ot.w);
    f3 gFo = mk3(0.f, 0.f, 0.f), gTo = gFo;
    /* a dead slot is copied through: its cotangent too */
    f3 gp = xyz(lam_out[3 * i]), gv = xyz(lam_out[3 * i + 1]), gw = xyz(lam_out[3 * i + 2]);
    if (pid >= 0) {
      const f3 p = xyz(pr), v = xyz(vm), w = xyz(ot);
      const float r = pr.w, m = vm.w;
      f3 F = cvt(P.gravity) * m, T = mk3(0.f, 0.f, 0.f);
      const int nc = min((int)st.nn[i], LIST_STRIDE);
      for (int k = 0; k < nc; ++k) {
        const int j = (int)st.list[(size_t)i * LIST_STRIDE + k];
        const float4 pb = st.pos[j], vb = st.vel[j], ob = st.omg[j];
        f3 Fc, Tc;
        pair_ref(P, p, r, v, m, w, ta, xyz(pb), pb.w, xyz(vb), vb.w, xyz(ob), packed_type(ob.w), Fc, Tc);
        F = F + Fc;
        if (P.rotation) T = T + Tc;
      }
      f3 Fw, Tw;
      wall_forces(P, ta, p, r, v, w, m, Fw, Tw, nullptr);
      /* integrate: v' = v + (Fw + F) / m dt, x' = x + v' dt; with torque,
       * w1 = w + Iinv (T + Tw) dt and w' = w1 - w1 rr */
      const float dt = P.dt;
      const f3 acc = (Fw + F) * (1.0f / m);
      const f3 v1 = v + acc * dt;
      const f3 p1 = p + v1 * dt;
      const f3 fl = T + Tw;
      const bool spin = P.rotation && dot(fl, fl) > 0.0f;
      const float iinv = P.per_particle ? 1.0f / (0.4f * r * r * m) : P.t_inertia[ta];
      const f3 w1 = w + (iinv * fl) * dt;
      const float4 lo0 = lam_out[3 * i], lo1 = lam_out[3 * i + 1], lo2 = lam_out[3 * i + 2];
      /* the loss reads the output position through the moments */
      f3 gp1 = xyz(lo0);
      gp1.x += m * (g.sx + 2.0f * g.sxx * p1.x + g.sxy * p1.y);
      gp1.y += m * (g.sy + 2.0f * g.syy * p1.y + g.sxy * p1.x);
      f3 gv1 = xyz(lo1);
      const f3 gw2 = xyz(lo2);
      gp = gp1;                      /* x' = x + v' dt */
      gv1 = gv1 + gp1 * dt;
      gv = gv1;                      /* v' = v + acc dt */
      gFo = gv1 * (dt / m);          /* acc = (Fw + F) / m */
      if (spin) {
        const f3 gw1 = gw2 - gw2 * P.roll_res;   /* w' = w1 - w1 rr */
        grr = -(double)dot(w1, gw2);
        gw = gw1;                                /* w1 = w + Iinv fl dt */
        gTo = gw1 * (iinv * dt);
      } else {
        gw = gw2;
      }
    }
    lamF[2 * i] = make_float4(gFo.x, gFo.y, gFo.z, 0.f);
    lamF[2 * i + 1] = make_float4(gTo.x, gTo.y, gTo.z, 0.f);
    lamIn[3 * i] = make_float4(gp.x, gp.y, gp.z, 0.f);
    lamIn[3 * i + 1] = make_float4(gv.x, gv.y, gv.z, 0.f);
    lamIn[3 * i + 2] = make_float4(gw.x, gw.y, gw.z, 0.f);
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int o = 32; o > 0; o >>= 1) grr += __shfl_down(grr, o, 64);
  if (lane == 0) red[wave] = grr;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int w = 0; w < ADJ_THREADS / 64; ++w) s += red[w];
    part_rr[blockIdx.x] += s;
  }
}

/* ------------------------------------------------------------------------ */
/* Reverse step, part 2 (per sorted position i): the pair law's cotangents  */
/* reach i as the owner of its contacts (with its own force cotangent) and  */
/* as the partner of each contact's owner (with the owner's), then the      */
/* walls'.  The contact set is symmetric, so i's list names every owner     */
/* that lists i.  The result goes to i's slot before this step's sort.      */
/* ------------------------------------------------------------------------ */
__global__ __launch_bounds__(ADJ_THREADS) void k_adj_back2(const DevParams* __restrict__ Pp, AdjStep st,
                                                           const float4* __restrict__ lamF,
                                                           const float4* __restrict__ lamIn,
                                                           float4* __restrict__ lam_prev,
                                                           double* __restrict__ part_pp) {
  const DevParams& P = *Pp;
  __shared__ double red[3][ADJ_THREADS / 64];
  const int i = blockIdx.x * ADJ_THREADS + threadIdx.x;
  float gKn = 0.f, gCn = 0.f, gFk = 0.f;
  if (i < st.n) {
    const float4 pr = st.pos[i], vm = st.vel[i], ot = st.omg[i];
    const int pid = packed_pid(ot.w), ta = packed_type(ot.w);
    f3 gp = xyz(lamIn[3 * i]), gv = xyz(lamIn[3 * i + 1]), gw = xyz(lamIn[3 * i + 2]);
    if (pid >= 0) {
      const f3 p = xyz(pr), v = xyz(vm), w = xyz(ot);
      const float r = pr.w, m = vm.w;
      const f3 gF = xyz(lamF[2 * i]), gT = xyz(lamF[2 * i + 1]);
      const int nc = min((int)st.nn[i], LIST_STRIDE);
      for (int k = 0; k < nc; ++k) {
        const int j = (int)st.list[(size_t)i * LIST_STRIDE + k];
        const float4 pb = st.pos[j], vb = st.vel[j], ob = st.omg[j];
        const f3 qb = xyz(pb), ub = xyz(vb), wb = xyz(ob);
        const int tb = packed_type(ob.w);
        f3 dpb = mk3(0.f, 0.f, 0.f), dvb = dpb, dwb = dpb;
        /* i owns the contact: Fc(i; j) with i's cotangent */
        pair_ref_vjp(P, p, r, v, m, w, ta, qb, pb.w, ub, vb.w, wb, tb, gF, gT, gp, gv, gw, dpb, dvb, dwb, gKn,
                     gCn, gFk);
        /* j owns the contact: Fc(j; i) with j's cotangent, i as partner */
        const f3 hF = xyz(lamF[2 * j]), hT = xyz(lamF[2 * j + 1]);
        f3 dpa = mk3(0.f, 0.f, 0.f), dva = dpa, dwa = dpa;
        float k0 = 0.f, k1 = 0.f, k2 = 0.f;
        pair_ref_vjp(P, qb, pb.w, ub, vb.w, wb, tb, p, r, v, m, w, ta, hF, hT, dpa, dva, dwa, gp, gv, gw, k0, k1,
                     k2);
      }
      wall_vjp(P, ta, p, r, v, w, m, gF, gT, gp, gv, gw);
    }
    const int prev = st.idx[i];
    lam_prev[3 * prev] = make_float4(gp.x, gp.y, gp.z, 0.f);
    lam_prev[3 * prev + 1] = make_float4(gv.x, gv.y, gv.z, 0.f);
    lam_prev[3 * prev + 2] = make_float4(gw.x, gw.y, gw.z, 0.f);
  }
  double s[3] = {gKn, gCn, gFk};
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    double x = s[k];
    for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o, 64);
    if (lane == 0) red[k][wave] = x;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    double x = 0.0;
    for (int w = 0; w < ADJ_THREADS / 64; ++w) x += red[threadIdx.x][w];
    part_pp[(size_t)blockIdx.x * 3 + threadIdx.x] += x;
  }
}

void launch_adj_moments(int n, const float4* pos, const float4* vel, const float4* omg, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_adj_moments, dim3((n + ADJ_THREADS - 1) / ADJ_THREADS), dim3(ADJ_THREADS), 0, s, n, pos, vel,
                     omg, out);
}

void launch_adj_back(const DevParams* dP, const AdjStep& st, const float4* lam_out, float4* lamF, float4* lamIn,
                     float4* lam_prev, const LossCoef& g, double* part_rr, double* part_pp, hipStream_t s) {
  const dim3 grid((st.n + ADJ_THREADS - 1) / ADJ_THREADS), blk(ADJ_THREADS);
  hipLaunchKernelGGL(k_adj_back1, grid, blk, 0, s, dP, st, lam_out, lamF, lamIn, g, part_rr);
  hipLaunchKernelGGL(k_adj_back2, grid, blk, 0, s, dP, st, lamF, lamIn, lam_prev, part_pp);
}

}  // namespace bdem
