// Force-kernel launches of one potential kind: the lpmdx_system member
// templates that pick the kernel variant (staged or not, lanes per atom,
// species reads, compile-time Sutton-Chen exponents).  Included by the
// engine_kinds_*.hip files, each of which instantiates launch_kind for a few
// kinds, so the kernel variants of different kinds compile in parallel.
#pragma once

#include "engine_system.hpp"

// Pair kinds: one force launch; EAM kinds: the rho pass, then the force pass.
template <int KIND, bool EVAL> void lpmdx_system::launch_kind(ForceArgs& A, long nb, double rc2_rho) {
  if constexpr (metal_kind(KIND)) {
    if (A.P.spc_sum >= 0)
      launch_eam_s<KIND, EVAL, true>(A, nb, rc2_rho);
    else
      launch_eam_s<KIND, EVAL, false>(A, nb, rc2_rho);
  } else {
    (void)rc2_rho;
    if (reads_species(A))
      launch_pair_s<KIND, EVAL, true>(A, nb);
    else
      launch_pair_s<KIND, EVAL, false>(A, nb);
  }
}

template <int KIND, bool EVAL, bool SPC> void lpmdx_system::launch_pair_s(ForceArgs& A, long nb) {
  hipEvent_t e0 = prof_begin();
  if (staged) {
    const int parts = SPC ? kStageS : 0;
    const BrickLaunch L{nbrick, brick_threads, stage_bytes(halo_cap, parts), st, bhalo.p, grid.brick, halo_cap};
    switch (lpa) {
    case 1: BrickKernels<KIND, 1, EVAL>::template pair<SPC>(L, A, parts); break;
    case 2: BrickKernels<KIND, 2, EVAL>::template pair<SPC>(L, A, parts); break;
    default: BrickKernels<KIND, 4, EVAL>::template pair<SPC>(L, A, parts); break;
    }
    prof_end(kProfForce, e0);
    return;
  }
  switch (lpa) {
  case 1: k_pair_force<KIND, 1, EVAL, SPC><<<nb, kBlock, 0, st>>>(A); break;
  case 2: k_pair_force<KIND, 2, EVAL, SPC><<<nb, kBlock, 0, st>>>(A); break;
  case 4: k_pair_force<KIND, 4, EVAL, SPC><<<nb, kBlock, 0, st>>>(A); break;
  case 8: k_pair_force<KIND, 8, EVAL, SPC><<<nb, kBlock, 0, st>>>(A); break;
  default: k_pair_force<KIND, 16, EVAL, SPC><<<nb, kBlock, 0, st>>>(A); break;
  }
  prof_end(kProfForce, e0);
}
// Unstaged Sutton-Chen steps with the common integer exponent pairs run
// kernels with the exponents fixed at compile time (bit-identical results,
// about 20 fewer instructions per neighbour).
template <int KIND, bool EVAL, bool SPC> bool lpmdx_system::launch_sc_fixed(ForceArgs& A, ForceArgs& R, long nb) {
  if constexpr (KIND != kSuttonChen || EVAL) {
    return false;
  } else {
    const int ex = sc_exp(A.P.ni, A.P.mi);
    if (lpa > 4 || A.P.ni <= 0 || A.P.mi <= 0)
      return false;
    if (ex == sc_exp(9, 6))
      return launch_eam_fixed<KIND, SPC, sc_exp(9, 6)>(A, R, nb);
    if (ex == sc_exp(12, 6))
      return launch_eam_fixed<KIND, SPC, sc_exp(12, 6)>(A, R, nb);
    return false;
  }
}
template <int KIND, bool SPC, int EXP> bool lpmdx_system::launch_eam_fixed(ForceArgs& A, ForceArgs& R, long nb) {
  hipEvent_t e0 = prof_begin();
  if (staged) {
    const int parts = kStageW | (SPC ? kStageS : 0);
    const BrickLaunch Lr{nbrick, brick_threads, stage_bytes(halo_cap, 0), st, bhalo.p, grid.brick, halo_cap};
    const BrickLaunch Lf{nbrick, brick_threads, stage_bytes(halo_cap, parts), st, bhalo.p, grid.brick, halo_cap};
    switch (lpa) {
    case 1: BrickKernels<KIND, 1, false>::template rho<EXP>(Lr, R, w_rho.p, rho.p); break;
    case 2: BrickKernels<KIND, 2, false>::template rho<EXP>(Lr, R, w_rho.p, rho.p); break;
    default: BrickKernels<KIND, 4, false>::template rho<EXP>(Lr, R, w_rho.p, rho.p); break;
    }
    hipEvent_t e1 = prof_next(kProfRho, e0);
    A.w_rho = w_rho.p;
    switch (lpa) {
    case 1: BrickKernels<KIND, 1, false>::template eam<SPC, EXP>(Lf, A, parts); break;
    case 2: BrickKernels<KIND, 2, false>::template eam<SPC, EXP>(Lf, A, parts); break;
    default: BrickKernels<KIND, 4, false>::template eam<SPC, EXP>(Lf, A, parts); break;
    }
    prof_end(kProfForce, e1, false);
    return true;
  }
  switch (lpa) {
  case 1: k_eam_rho<KIND, 1, false, EXP><<<nb, kBlock, 0, st>>>(R, w_rho.p, rho.p); break;
  case 2: k_eam_rho<KIND, 2, false, EXP><<<nb, kBlock, 0, st>>>(R, w_rho.p, rho.p); break;
  default: k_eam_rho<KIND, 4, false, EXP><<<nb, kBlock, 0, st>>>(R, w_rho.p, rho.p); break;
  }
  hipEvent_t e1 = prof_next(kProfRho, e0);
  A.w_rho = w_rho.p;
  switch (lpa) {
  case 1: k_eam_force<KIND, 1, false, SPC, EXP><<<nb, kBlock, 0, st>>>(A); break;
  case 2: k_eam_force<KIND, 2, false, SPC, EXP><<<nb, kBlock, 0, st>>>(A); break;
  default: k_eam_force<KIND, 4, false, SPC, EXP><<<nb, kBlock, 0, st>>>(A); break;
  }
  prof_end(kProfForce, e1, false);
  return true;
}
template <int KIND, bool EVAL, bool SPC> void lpmdx_system::launch_eam_s(ForceArgs& A, long nb, double rc2_rho) {
  ForceArgs R = A;
  R.P.rc2 = rc2_rho;
  R.partials = partials2.p;
  if (launch_sc_fixed<KIND, EVAL, SPC>(A, R, nb))
    return;
  hipEvent_t e0 = prof_begin();
  if (staged) {
    const int parts = kStageW | (SPC ? kStageS : 0);
    const BrickLaunch Lr{nbrick, brick_threads, stage_bytes(halo_cap, 0), st, bhalo.p, grid.brick, halo_cap};
    const BrickLaunch Lf{nbrick, brick_threads, stage_bytes(halo_cap, parts), st, bhalo.p, grid.brick, halo_cap};
    switch (lpa) {
    case 1: BrickKernels<KIND, 1, EVAL>::rho(Lr, R, w_rho.p, rho.p); break;
    case 2: BrickKernels<KIND, 2, EVAL>::rho(Lr, R, w_rho.p, rho.p); break;
    default: BrickKernels<KIND, 4, EVAL>::rho(Lr, R, w_rho.p, rho.p); break;
    }
    hipEvent_t e1 = prof_next(kProfRho, e0);
    A.w_rho = w_rho.p;
    switch (lpa) {
    case 1: BrickKernels<KIND, 1, EVAL>::template eam<SPC>(Lf, A, parts); break;
    case 2: BrickKernels<KIND, 2, EVAL>::template eam<SPC>(Lf, A, parts); break;
    default: BrickKernels<KIND, 4, EVAL>::template eam<SPC>(Lf, A, parts); break;
    }
    prof_end(kProfForce, e1, false);
    return;
  }
  switch (lpa) {
  case 1: k_eam_rho<KIND, 1, EVAL><<<nb, kBlock, 0, st>>>(R, w_rho.p, rho.p); break;
  case 2: k_eam_rho<KIND, 2, EVAL><<<nb, kBlock, 0, st>>>(R, w_rho.p, rho.p); break;
  case 4: k_eam_rho<KIND, 4, EVAL><<<nb, kBlock, 0, st>>>(R, w_rho.p, rho.p); break;
  case 8: k_eam_rho<KIND, 8, EVAL><<<nb, kBlock, 0, st>>>(R, w_rho.p, rho.p); break;
  default: k_eam_rho<KIND, 16, EVAL><<<nb, kBlock, 0, st>>>(R, w_rho.p, rho.p); break;
  }
  hipEvent_t e1 = prof_next(kProfRho, e0);
  A.w_rho = w_rho.p;
  switch (lpa) {
  case 1: k_eam_force<KIND, 1, EVAL, SPC><<<nb, kBlock, 0, st>>>(A); break;
  case 2: k_eam_force<KIND, 2, EVAL, SPC><<<nb, kBlock, 0, st>>>(A); break;
  case 4: k_eam_force<KIND, 4, EVAL, SPC><<<nb, kBlock, 0, st>>>(A); break;
  case 8: k_eam_force<KIND, 8, EVAL, SPC><<<nb, kBlock, 0, st>>>(A); break;
  default: k_eam_force<KIND, 16, EVAL, SPC><<<nb, kBlock, 0, st>>>(A); break;
  }
  prof_end(kProfForce, e1, false);
}

// one potential kind's launchers, both with and without energies
#define LPMDX_INSTANTIATE_KIND(K) \
  template void lpmdx_system::launch_kind<K, true>(ForceArgs&, long, double); \
  template void lpmdx_system::launch_kind<K, false>(ForceArgs&, long, double);
