// qmc_basis.hpp -- how the orbital and Jastrow loops walk the shells and
// nuclei of a system.
//
// RtBasis reads the shell table (kind, centre, primitive count, offsets)
// from DevSys at run time: it serves any system, but every shell costs a
// scalar load that the wave has to wait for before it can go on.
//
// CtBasis<Bd> takes the same table from a compile-time descriptor Bd, so
// the shell and primitive loops unroll completely and every parameter
// address becomes a constant offset into DevSys: the compiler batches and
// hoists the scalar loads instead of waiting once per primitive.  The
// parameter VALUES (exponents, coefficients, MO matrix, nuclear positions,
// Jastrow) still come from DevSys, so one specialisation serves every
// geometry and parameter set with that shell structure.  The host selects
// a specialisation only when the system's shell table equals Bd's exactly
// (basis_matches).
#pragma once

#include <type_traits>

#include "qmc_sys.hpp"

namespace qmc {

template <int I, int N, class F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (I < N) {
    f(std::integral_constant<int, I>{});
    static_for<I + 1, N>(f);
  }
}

constexpr int shell_nao(int kind) { return kind == SH_SP ? 4 : (kind == SH_P ? 3 : 1); }

template <class T>
struct RtShell {
  const DevSys<T>& S;
  int i;
  __device__ __forceinline__ int kind() const { return S.sh_kind[i]; }
  __device__ __forceinline__ int center() const { return S.sh_center[i]; }
  __device__ __forceinline__ int nprim() const { return S.sh_nprim[i]; }
  __device__ __forceinline__ int p0() const { return S.sh_prim0[i]; }
  __device__ __forceinline__ int a0() const { return S.sh_ao0[i]; }
  __device__ __forceinline__ int m() const { return S.sh_m[i]; }
  __device__ __forceinline__ int pw(int k) const { return S.prim_pow[S.sh_prim0[i] + k]; }
};

struct RtBasis {
  static constexpr bool lds_mo = false;  // MO coefficients read from DevSys
  static constexpr int n_nuc_ct = 0;  // nucleus count unknown at compile time
  template <class T, class F>
  static __device__ __forceinline__ void shells(const DevSys<T>& S, F&& f) {
    for (int i = 0; i < S.n_shell; ++i) f(RtShell<T>{S, i});
  }
  template <class T, class F>
  static __device__ __forceinline__ void nuclei(const DevSys<T>& S, F&& f) {
    for (int I = 0; I < S.n_nuc; ++I) f(I);
  }
};

// offsets of shell i in a compile-time descriptor
template <class Bd>
constexpr int bd_p0(int i) {
  int s = 0;
  for (int j = 0; j < i; ++j) s += Bd::nprim[j];
  return s;
}
template <class Bd>
constexpr int bd_a0(int i) {
  int s = 0;
  for (int j = 0; j < i; ++j) s += shell_nao(Bd::kind[j]);
  return s;
}

template <class Bd, int I>
struct CtShell {
  static constexpr int kind() { return Bd::kind[I]; }
  static constexpr int center() { return Bd::center[I]; }
  static constexpr int nprim() { return Bd::nprim[I]; }
  static constexpr int p0() { return bd_p0<Bd>(I); }
  static constexpr int a0() { return bd_a0<Bd>(I); }
  static constexpr int m() { return Bd::m[I]; }
  static constexpr int pw(int k) { return Bd::pw[bd_p0<Bd>(I) + k]; }
};

template <class Bd>
struct CtBasis {
  static constexpr bool lds_mo = false;
  static constexpr int n_nuc_ct = Bd::n_nuc;
  template <class T, class F>
  static __device__ __forceinline__ void shells(const DevSys<T>&, F&& f) {
    static_for<0, Bd::n_shell>([&](auto I) { f(CtShell<Bd, decltype(I)::value>{}); });
  }
  template <class T, class F>
  static __device__ __forceinline__ void nuclei(const DevSys<T>&, F&& f) {
    static_for<0, Bd::n_nuc>([&](auto I) { f((int)decltype(I)::value); });
  }
};

// ---------------------------------------------------------------------------
// Compiled shell structures.  Each lists, per shell in DevSys order, the
// kind, centre, primitive count and P1 component, and per primitive the STO
// power (0 for Gaussians).  They follow the AO order of the constructors.
// ---------------------------------------------------------------------------

// H2MoleculeMO: Slater1s on each proton (hydrogen.rs:28-35)
struct BdH2 {
  static constexpr const char* name = "h2-slater1s";
  static constexpr int nu = 1, nd = 1, n_nuc = 2, n_shell = 2;
  static constexpr int kind[] = {SH_SLATER1S, SH_SLATER1S};
  static constexpr int center[] = {0, 1};
  static constexpr int nprim[] = {1, 1};
  static constexpr int m[] = {0, 0};
  static constexpr int pw[] = {0, 0};
};

// Li, three STO orbitals of 3 terms r^0, r^1, r^2 (the seeded Li atom)
struct BdLiSeeded {
  static constexpr const char* name = "li-sto3term";
  static constexpr int nu = 2, nd = 1, n_nuc = 1, n_shell = 3;
  static constexpr int kind[] = {SH_STO, SH_STO, SH_STO};
  static constexpr int center[] = {0, 0, 0};
  static constexpr int nprim[] = {3, 3, 3};
  static constexpr int m[] = {0, 0, 0};
  static constexpr int pw[] = {0, 1, 2, 0, 1, 2, 0, 1, 2};
};

// Li, three init_li_sto orbitals of 7 terms (slater.rs:68-89)
struct BdLiRef {
  static constexpr const char* name = "li-init-li-sto";
  static constexpr int nu = 2, nd = 1, n_nuc = 1, n_shell = 3;
  static constexpr int kind[] = {SH_STO, SH_STO, SH_STO};
  static constexpr int center[] = {0, 0, 0};
  static constexpr int nprim[] = {7, 7, 7};
  static constexpr int m[] = {0, 0, 0};
  static constexpr int pw[] = {1, 1, 2, 2, 2, 2, 3, 1, 1, 2, 2, 2, 2, 3, 1, 1, 2, 2, 2, 2, 3};
};

// CH4 STO-3G: C 1s, C 2sp, H 1s x 4, three primitives each
struct BdCH4Sto3g {
  static constexpr const char* name = "ch4-sto3g";
  static constexpr int nu = 5, nd = 5, n_nuc = 5, n_shell = 6;
  static constexpr int kind[] = {SH_S, SH_SP, SH_S, SH_S, SH_S, SH_S};
  static constexpr int center[] = {0, 0, 1, 2, 3, 4};
  static constexpr int nprim[] = {3, 3, 3, 3, 3, 3};
  static constexpr int m[] = {0, 0, 0, 0, 0, 0};
  static constexpr int pw[] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
};

// CH4 6-31G (methane.rs:697-776): C 1s(6), 2sp(3), 3sp(1); H 1s(3) x 4; H 2s(1) x 4
struct BdCH4631g {
  static constexpr const char* name = "ch4-631g";
  static constexpr int nu = 5, nd = 5, n_nuc = 5, n_shell = 11;
  static constexpr int kind[] = {SH_S, SH_SP, SH_SP, SH_S, SH_S, SH_S, SH_S, SH_S, SH_S, SH_S, SH_S};
  static constexpr int center[] = {0, 0, 0, 1, 2, 3, 4, 1, 2, 3, 4};
  static constexpr int nprim[] = {6, 3, 1, 3, 3, 3, 3, 1, 1, 1, 1};
  static constexpr int m[] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  static constexpr int pw[] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0,
                               0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
};

// true when the system's shell table is exactly Bd's
template <class Bd, class T>
bool basis_matches(const DevSys<T>& D) {
  if (D.n_up != Bd::nu || D.n_dn != Bd::nd || D.n_nuc != Bd::n_nuc || D.n_shell != Bd::n_shell) return false;
  int np = 0, na = 0;
  for (int i = 0; i < Bd::n_shell; ++i) {
    if (D.sh_kind[i] != Bd::kind[i] || D.sh_center[i] != Bd::center[i] || D.sh_nprim[i] != Bd::nprim[i] ||
        D.sh_prim0[i] != np || D.sh_ao0[i] != na)
      return false;
    if (Bd::kind[i] == SH_P1 && D.sh_m[i] != Bd::m[i]) return false;
    for (int k = 0; k < Bd::nprim[i]; ++k)
      if ((Bd::kind[i] == SH_STO || Bd::kind[i] == SH_SLATER1S) && D.prim_pow[np + k] != Bd::pw[np + k])
        return false;
    np += Bd::nprim[i];
    na += shell_nao(Bd::kind[i]);
  }
  return na == D.n_ao;
}

// CtBasis whose MO coefficients come from a per-workgroup LDS table
// (lds_mo_tab, filled at kernel start) instead of DevSys: the compiler then
// keeps them in vector registers rather than in scalar registers that it
// spills to VGPR lanes and restores with v_readlane.  Used by the FP32
// single-electron sweep kernels.
template <class Bd>
struct CtBasisL : CtBasis<Bd> {
  static constexpr bool lds_mo = true;
  static constexpr int n_ao = bd_a0<Bd>(Bd::n_shell);  // the table must hold them
};

}  // namespace qmc
