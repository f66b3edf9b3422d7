// qmc_io.cpp -- system descriptions as JSON text (replaces the YAML loaders
// of io/config.rs:20-35, which deserialize H2 only).  A small
// recursive-descent JSON reader turns a document into the public
// constructors: the reference's serde layout of H2MoleculeMO
// (hydrogen.rs:27-35), named molecules, or the generic description that
// mirrors qmc_system_desc.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "qmc.h"

int qmc_io_fail(int code, const char* msg);  // sets qmc_last_error (qmc_host.cpp)

namespace {

struct JVal {
  enum Kind { NUL, BOOL, NUM, STR, ARR, OBJ } k = NUL;
  double num = 0.0;
  bool b = false;
  std::string str;
  std::vector<JVal> arr;
  std::vector<std::pair<std::string, JVal>> obj;
  const JVal* get(const char* key) const {
    if (k != OBJ) return nullptr;
    for (const auto& kv : obj)
      if (kv.first == key) return &kv.second;
    return nullptr;
  }
};

constexpr int MAX_DEPTH = 64;

struct Reader {
  const char* p;
  const char* end;
  std::string err;

  void ws() {
    while (p < end && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) ++p;
  }
  bool fail(const char* what) {
    if (err.empty()) err = std::string(what) + " at offset " + std::to_string(off0 ? (long)(p - off0) : 0L);
    return false;
  }
  const char* off0 = nullptr;

  bool lit(const char* s) {
    const size_t n = std::strlen(s);
    if ((size_t)(end - p) < n || std::strncmp(p, s, n) != 0) return fail("bad literal");
    p += n;
    return true;
  }
  bool string(std::string& out) {
    if (p >= end || *p != '"') return fail("expected string");
    ++p;
    while (p < end && *p != '"') {
      char c = *p++;
      if (c == '\\') {
        if (p >= end) return fail("bad escape");
        const char e = *p++;
        switch (e) {
          case '"': case '\\': case '/': out += e; break;
          case 'b': out += '\b'; break;
          case 'f': out += '\f'; break;
          case 'n': out += '\n'; break;
          case 'r': out += '\r'; break;
          case 't': out += '\t'; break;
          case 'u': {
            if (end - p < 4) return fail("bad \\u escape");
            unsigned v = 0;
            for (int i = 0; i < 4; ++i) {
              const char h = *p++;
              v = v * 16 + (h >= '0' && h <= '9' ? h - '0' : h >= 'a' && h <= 'f' ? h - 'a' + 10
                                                             : h >= 'A' && h <= 'F' ? h - 'A' + 10 : 99);
              if (v > 0xFFFF) return fail("bad \\u escape");
            }
            out += v < 0x80 ? (char)v : '?';  // keys and names are ASCII
            break;
          }
          default: return fail("bad escape");
        }
      } else {
        out += c;
      }
    }
    if (p >= end) return fail("unterminated string");
    ++p;
    return true;
  }
  bool number(double& v) {
    const char* s = p;
    if (p < end && (*p == '-' || *p == '+')) ++p;
    while (p < end && ((*p >= '0' && *p <= '9') || *p == '.' || *p == 'e' || *p == 'E' || *p == '-' || *p == '+'))
      ++p;
    std::string tok(s, p);
    char* q = nullptr;
    v = std::strtod(tok.c_str(), &q);
    if (tok.empty() || q != tok.c_str() + tok.size()) return fail("bad number");
    return true;
  }
  bool value(JVal& v, int depth) {
    if (depth > MAX_DEPTH) return fail("nesting too deep");
    ws();
    if (p >= end) return fail("unexpected end");
    const char c = *p;
    if (c == '{') {
      v.k = JVal::OBJ;
      ++p;
      ws();
      if (p < end && *p == '}') {
        ++p;
        return true;
      }
      for (;;) {
        ws();
        std::pair<std::string, JVal> kv;
        if (!string(kv.first)) return false;
        ws();
        if (p >= end || *p != ':') return fail("expected ':'");
        ++p;
        if (!value(kv.second, depth + 1)) return false;
        v.obj.push_back(std::move(kv));
        ws();
        if (p < end && *p == ',') {
          ++p;
          continue;
        }
        if (p < end && *p == '}') {
          ++p;
          return true;
        }
        return fail("expected ',' or '}'");
      }
    }
    if (c == '[') {
      v.k = JVal::ARR;
      ++p;
      ws();
      if (p < end && *p == ']') {
        ++p;
        return true;
      }
      for (;;) {
        JVal e;
        if (!value(e, depth + 1)) return false;
        v.arr.push_back(std::move(e));
        ws();
        if (p < end && *p == ',') {
          ++p;
          continue;
        }
        if (p < end && *p == ']') {
          ++p;
          return true;
        }
        return fail("expected ',' or ']'");
      }
    }
    if (c == '"') {
      v.k = JVal::STR;
      return string(v.str);
    }
    if (c == 't') {
      v.k = JVal::BOOL;
      v.b = true;
      return lit("true");
    }
    if (c == 'f') {
      v.k = JVal::BOOL;
      return lit("false");
    }
    if (c == 'n') return lit("null");
    v.k = JVal::NUM;
    return number(v.num);
  }
};

// ---- typed access with messages ------------------------------------------
struct Conv {
  std::string err;
  bool bad(const std::string& m) {
    if (err.empty()) err = m;
    return false;
  }
  bool num(const JVal* v, const char* what, double& out) {
    if (!v || v->k != JVal::NUM) return bad(std::string(what) + ": expected a number");
    out = v->num;
    return true;
  }
  bool num_or(const JVal* obj, const char* key, double dflt, double& out) {
    const JVal* v = obj ? obj->get(key) : nullptr;
    if (!v) {
      out = dflt;
      return true;
    }
    return num(v, key, out);
  }
  bool integer(const JVal* v, const char* what, int64_t& out) {
    double d;
    if (!num(v, what, d)) return false;
    if (d != std::floor(d) || std::fabs(d) > 9.0e15) return bad(std::string(what) + ": expected an integer");
    out = (int64_t)d;
    return true;
  }
  bool nums(const JVal* v, const char* what, std::vector<double>& out) {
    if (!v || v->k != JVal::ARR) return bad(std::string(what) + ": expected an array of numbers");
    for (const auto& e : v->arr) {
      double d;
      if (!num(&e, what, d)) return false;
      out.push_back(d);
    }
    return true;
  }
  bool ints(const JVal* v, const char* what, std::vector<int32_t>& out) {
    if (!v || v->k != JVal::ARR) return bad(std::string(what) + ": expected an array of integers");
    for (const auto& e : v->arr) {
      int64_t i;
      if (!integer(&e, what, i)) return false;
      if (i < INT32_MIN || i > INT32_MAX) return bad(std::string(what) + ": integer out of range");
      out.push_back((int32_t)i);
    }
    return true;
  }
  // [[a, b, ...], ...] with rows of equal length `cols` (0: any, returned)
  bool matrix(const JVal* v, const char* what, int& cols, std::vector<double>& out) {
    if (!v || v->k != JVal::ARR) return bad(std::string(what) + ": expected an array of rows");
    for (const auto& row : v->arr) {
      std::vector<double> r;
      if (!nums(&row, what, r)) return false;
      if (cols == 0) cols = (int)r.size();
      if ((int)r.size() != cols || cols == 0) return bad(std::string(what) + ": ragged or empty rows");
      out.insert(out.end(), r.begin(), r.end());
    }
    return true;
  }
  bool vec3(const JVal* v, const char* what, double out[3]) {
    std::vector<double> r;
    if (!nums(v, what, r)) return false;
    if (r.size() != 3) return bad(std::string(what) + ": expected [x, y, z]");
    for (int c = 0; c < 3; ++c) out[c] = r[c];
    return true;
  }
  bool flags(const JVal* v, uint32_t dflt, uint32_t& out) {
    if (!v) {
      out = dflt;
      return true;
    }
    if (v->k == JVal::NUM) {
      int64_t i;
      if (!integer(v, "flags", i) || i < 0 || i > 0xFFFF) return bad("flags: expected a small non-negative integer");
      out = (uint32_t)i;
      return true;
    }
    if (v->k != JVal::ARR) return bad("flags: expected an integer or a list of names");
    out = 0;
    for (const auto& e : v->arr) {
      if (e.k != JVal::STR) return bad("flags: expected names");
      if (e.str == "j2_grad_quirk")
        out |= QMC_FLAG_J2_GRAD_QUIRK;
      else if (e.str == "j2_lap_quirk")
        out |= QMC_FLAG_J2_LAP_QUIRK;
      else if (e.str == "hf_nn_refsign")
        out |= QMC_FLAG_HF_NN_REFSIGN;
      else if (e.str == "singular_identity")
        out |= QMC_FLAG_SINGULAR_IDENTITY;
      else if (e.str == "reference")
        out |= QMC_FLAGS_REFERENCE;
      else
        return bad("flags: unknown flag '" + e.str + "'");
    }
    return true;
  }
};

int finish(int rc, qmc_system* sys, qmc_system** out, uint32_t flags, bool set_flags) {
  if (rc) return rc;
  if (set_flags && (rc = qmc_set_flags(sys, flags))) {
    qmc_system_destroy(sys);
    return rc;
  }
  *out = sys;
  return QMC_OK;
}

// H2MoleculeMO {orbital1: Slater1s{alpha, center}, orbital2, jastrow:
// Jastrow1{cusp_param}} (hydrogen.rs:27-35), or the keys H1/H2 {alpha, R}
// and J {F} of conf/h2.yaml
int h2_layout(const JVal& root, int32_t precision, qmc_system** out, Conv& cv) {
  const bool ref = root.get("orbital1") != nullptr;
  const JVal* o[2] = {root.get(ref ? "orbital1" : "H1"), root.get(ref ? "orbital2" : "H2")};
  const JVal* j = root.get(ref ? "jastrow" : "J");
  double alpha[2], c[2][3], F, sigma;
  for (int i = 0; i < 2; ++i) {
    if (!o[i] || o[i]->k != JVal::OBJ) return qmc_io_fail(QMC_ERR_INVALID, "H2: both orbitals are needed");
    if (!cv.num(o[i]->get("alpha"), "alpha", alpha[i]) || !cv.vec3(o[i]->get(ref ? "center" : "R"), "center", c[i]))
      return qmc_io_fail(QMC_ERR_INVALID, cv.err.c_str());
  }
  if (!j || !cv.num(j->get(ref ? "cusp_param" : "F"), "cusp_param", F) ||
      !cv.num_or(&root, "init_sigma", 1.0, sigma))
    return qmc_io_fail(QMC_ERR_INVALID, cv.err.empty() ? "H2: jastrow.cusp_param is needed" : cv.err.c_str());
  uint32_t flags;
  if (!cv.flags(root.get("flags"), QMC_FLAG_HF_NN_REFSIGN, flags)) return qmc_io_fail(QMC_ERR_INVALID, cv.err.c_str());
  // two Slater1s AOs (one exponent each), psi = (phi1 + phi2)(r1) (phi1 + phi2)(r2) J
  const int32_t spins[2] = {1, -1}, init_nuc[2] = {0, 1};
  const double nuc[6] = {c[0][0], c[0][1], c[0][2], c[1][0], c[1][1], c[1][2]}, Z[2] = {1.0, 1.0};
  const double pexp[2] = {alpha[0], alpha[1]}, pcoef[2] = {1.0, 1.0}, mo[2] = {1.0, 1.0};
  qmc_ao aos[2] = {{QMC_AO_SLATER1S, 0, 0, 0, 1, 0}, {QMC_AO_SLATER1S, 1, 0, 0, 1, 1}};
  qmc_system_desc d{};
  d.n_elec = 2;
  d.spins = spins;
  d.n_nuc = 2;
  d.nuc_pos = nuc;
  d.nuc_charge = Z;
  d.n_ao = 2;
  d.aos = aos;
  d.prim_exp = pexp;
  d.prim_coef = pcoef;
  d.n_mo = 1;
  d.mo_up = mo;
  d.mo_dn = mo;
  d.jastrow_kind = QMC_JAS_PAIR;
  d.jas_F = F;
  d.init_nuc = init_nuc;
  d.init_sigma = sigma;
  d.flags = flags;
  return qmc_system_create(&d, precision, out);
}

int named(const JVal& root, const std::string& name, int32_t precision, qmc_system** out, Conv& cv) {
  qmc_system* s = nullptr;
  uint32_t flags = 0;
  const bool has_flags = root.get("flags") != nullptr;
  if (has_flags && !cv.flags(root.get("flags"), 0, flags)) return qmc_io_fail(QMC_ERR_INVALID, cv.err.c_str());
  if (name == "MethaneGTO") {
    double b_ee, b_en, ch;
    int64_t seed = 0;
    std::string basis = "6-31G";
    if (const JVal* b = root.get("basis")) {
      if (b->k != JVal::STR) return qmc_io_fail(QMC_ERR_INVALID, "basis: expected \"6-31G\" or \"STO-3G\"");
      basis = b->str;
    }
    if (!cv.num_or(&root, "b_ee", 1.5, b_ee) || !cv.num_or(&root, "b_en", 2.0, b_en) ||
        !cv.num_or(&root, "ch_bond", 2.05, ch) || (root.get("mo_seed") && !cv.integer(root.get("mo_seed"), "mo_seed", seed)))
      return qmc_io_fail(QMC_ERR_INVALID, cv.err.c_str());
    const int32_t bk = basis == "6-31G" ? QMC_BASIS_631G : basis == "STO-3G" ? QMC_BASIS_STO3G : -1;
    if (bk < 0) return qmc_io_fail(QMC_ERR_INVALID, "basis: expected \"6-31G\" or \"STO-3G\"");
    return finish(qmc_methane_gto_create(b_ee, b_en, ch, bk, (uint64_t)seed, precision, &s), s, out, flags, has_flags);
  }
  if (name == "BenzeneGTO") {
    double b_ee, b_en, cc, ch;
    if (!cv.num_or(&root, "b_ee", 1.5, b_ee) || !cv.num_or(&root, "b_en", 2.0, b_en) ||
        !cv.num_or(&root, "cc_bond", 2.64, cc) || !cv.num_or(&root, "ch_bond", 2.05, ch))
      return qmc_io_fail(QMC_ERR_INVALID, cv.err.c_str());
    return finish(qmc_benzene_gto_create(b_ee, b_en, cc, ch, precision, &s), s, out, flags, has_flags);
  }
  if (name == "LithiumSeeded") {
    int64_t seed = 0;
    if (root.get("seed") && !cv.integer(root.get("seed"), "seed", seed))
      return qmc_io_fail(QMC_ERR_INVALID, cv.err.c_str());
    return qmc_lithium_seeded_create((uint64_t)seed, has_flags ? flags : QMC_FLAGS_REFERENCE, precision, out);
  }
  if (name == "H2MoleculeMO") return h2_layout(root, precision, out, cv);
  return qmc_io_fail(QMC_ERR_INVALID, ("unknown system '" + name + "'").c_str());
}

int generic(const JVal& root, int32_t precision, qmc_system** out, Conv& cv) {
  std::vector<double> nuc, Z, pexp, pcoef, mo_up, mo_dn, c_anti, c_para, d_en;
  std::vector<int32_t> spins, ppow, init_nuc, species;
  int c3 = 3, nao_cols = 0, nao_cols_dn = 0, k_en_cols = 0;
  if (!cv.matrix(root.get("nuclei"), "nuclei", c3, nuc) || !cv.nums(root.get("charges"), "charges", Z) ||
      !cv.ints(root.get("spins"), "spins", spins))
    return qmc_io_fail(QMC_ERR_INVALID, cv.err.c_str());
  if (nuc.size() != 3 * Z.size()) return qmc_io_fail(QMC_ERR_INVALID, "nuclei and charges differ in length");
  const JVal* aos_v = root.get("aos");
  if (!aos_v || aos_v->k != JVal::ARR || aos_v->arr.empty())
    return qmc_io_fail(QMC_ERR_INVALID, "aos: expected a non-empty array");
  std::vector<qmc_ao> aos;
  for (const auto& a : aos_v->arr) {
    if (a.k != JVal::OBJ) return qmc_io_fail(QMC_ERR_INVALID, "aos: expected objects");
    qmc_ao ao{};
    const JVal* kind = a.get("kind");
    const std::string k = kind && kind->k == JVal::STR ? kind->str : "";
    ao.kind = k == "gto" ? QMC_AO_GTO : k == "sto" ? QMC_AO_STO : k == "slater1s" ? QMC_AO_SLATER1S : -1;
    if (ao.kind < 0) return qmc_io_fail(QMC_ERR_INVALID, "aos.kind: expected \"gto\", \"sto\" or \"slater1s\"");
    int64_t center = 0, l = 0, m = 0;
    if (!cv.integer(a.get("center"), "aos.center", center) || (a.get("l") && !cv.integer(a.get("l"), "aos.l", l)) ||
        (a.get("m") && !cv.integer(a.get("m"), "aos.m", m)))
      return qmc_io_fail(QMC_ERR_INVALID, cv.err.c_str());
    std::vector<double> e, c;
    std::vector<int32_t> pw;
    if (!cv.nums(a.get("exp"), "aos.exp", e) || !cv.nums(a.get("coef"), "aos.coef", c) ||
        (a.get("pow") && !cv.ints(a.get("pow"), "aos.pow", pw)))
      return qmc_io_fail(QMC_ERR_INVALID, cv.err.c_str());
    if (e.empty() || e.size() != c.size() || (!pw.empty() && pw.size() != e.size()))
      return qmc_io_fail(QMC_ERR_INVALID, "aos: exp, coef (and pow) must have the same non-zero length");
    for (int32_t q : pw)
      if (q < 0) return qmc_io_fail(QMC_ERR_INVALID, "aos.pow: powers must be >= 0");
    if (center < 0 || center >= (int64_t)Z.size() || l < 0 || l > 1 || m < 0 || m > 2)
      return qmc_io_fail(QMC_ERR_INVALID, "aos: centre, l or m out of range");
    ao.center = (int32_t)center;
    ao.l = (int32_t)l;
    ao.m = (int32_t)m;
    ao.nprim = (int32_t)e.size();
    ao.prim0 = (int32_t)pexp.size();
    pexp.insert(pexp.end(), e.begin(), e.end());
    pcoef.insert(pcoef.end(), c.begin(), c.end());
    if (pw.empty()) pw.assign(e.size(), 0);
    ppow.insert(ppow.end(), pw.begin(), pw.end());
    aos.push_back(ao);
  }
  nao_cols = (int)aos.size();
  if (!cv.matrix(root.get("mo_up"), "mo_up", nao_cols, mo_up)) return qmc_io_fail(QMC_ERR_INVALID, cv.err.c_str());
  if (const JVal* dn = root.get("mo_dn")) {
    nao_cols_dn = (int)aos.size();
    if (!cv.matrix(dn, "mo_dn", nao_cols_dn, mo_dn)) return qmc_io_fail(QMC_ERR_INVALID, cv.err.c_str());
    if (mo_dn.size() != mo_up.size()) return qmc_io_fail(QMC_ERR_INVALID, "mo_dn must have mo_up's shape");
  }
  qmc_system_desc d{};
  d.jastrow_kind = QMC_JAS_NONE;
  d.jas_F = 1.0;
  if (const JVal* j = root.get("jastrow")) {
    if (j->k != JVal::OBJ) return qmc_io_fail(QMC_ERR_INVALID, "jastrow: expected an object");
    const JVal* kind = j->get("kind");
    const std::string k = kind && kind->k == JVal::STR ? kind->str : "";
    d.jastrow_kind = k == "none" ? QMC_JAS_NONE : k == "pair" ? QMC_JAS_PAIR : k == "pade" ? QMC_JAS_PADE : -1;
    if (d.jastrow_kind < 0) return qmc_io_fail(QMC_ERR_INVALID, "jastrow.kind: expected \"none\", \"pair\" or \"pade\"");
    if (!cv.num_or(j, "F", 1.0, d.jas_F) || !cv.num_or(j, "b_ee", 0.0, d.b_ee) || !cv.num_or(j, "b_en", 0.0, d.b_en) ||
        !cv.num_or(j, "a_anti", 0.0, d.a_anti) || !cv.num_or(j, "a_para", 0.0, d.a_para))
      return qmc_io_fail(QMC_ERR_INVALID, cv.err.c_str());
    if (j->get("c_anti") || j->get("c_para")) {
      if (!cv.nums(j->get("c_anti"), "jastrow.c_anti", c_anti) || !cv.nums(j->get("c_para"), "jastrow.c_para", c_para))
        return qmc_io_fail(QMC_ERR_INVALID, cv.err.c_str());
      if (c_anti.size() != c_para.size()) return qmc_io_fail(QMC_ERR_INVALID, "c_anti and c_para differ in length");
    }
    if (j->get("d_en")) {
      if (!cv.matrix(j->get("d_en"), "jastrow.d_en", k_en_cols, d_en) ||
          !cv.ints(j->get("species"), "jastrow.species", species))
        return qmc_io_fail(QMC_ERR_INVALID, cv.err.c_str());
    }
  }
  if (root.get("init_nuc") && !cv.ints(root.get("init_nuc"), "init_nuc", init_nuc))
    return qmc_io_fail(QMC_ERR_INVALID, cv.err.c_str());
  if (!init_nuc.empty() && init_nuc.size() != spins.size())
    return qmc_io_fail(QMC_ERR_INVALID, "init_nuc must have one entry per electron");
  uint32_t flags;
  if (!cv.num_or(&root, "init_sigma", 1.0, d.init_sigma) || !cv.flags(root.get("flags"), 0, flags))
    return qmc_io_fail(QMC_ERR_INVALID, cv.err.c_str());
  d.n_elec = (int32_t)spins.size();
  d.spins = spins.data();
  d.n_nuc = (int32_t)Z.size();
  d.nuc_pos = nuc.data();
  d.nuc_charge = Z.data();
  d.n_ao = (int32_t)aos.size();
  d.aos = aos.data();
  d.prim_exp = pexp.data();
  d.prim_coef = pcoef.data();
  d.prim_pow = ppow.data();
  d.n_mo = (int32_t)(mo_up.size() / aos.size());
  d.mo_up = mo_up.data();
  d.mo_dn = mo_dn.empty() ? nullptr : mo_dn.data();
  d.k_ee = (int32_t)c_anti.size();
  d.c_anti = c_anti.empty() ? nullptr : c_anti.data();
  d.c_para = c_para.empty() ? nullptr : c_para.data();
  d.k_en = k_en_cols;
  d.n_species = k_en_cols ? (int32_t)(d_en.size() / k_en_cols) : 0;
  d.d_en = d_en.empty() ? nullptr : d_en.data();
  if (d.k_en > 0) {
    if (species.size() != Z.size()) return qmc_io_fail(QMC_ERR_INVALID, "jastrow.species: one entry per nucleus");
    for (int32_t sp : species)
      if (sp < 0 || sp >= d.n_species) return qmc_io_fail(QMC_ERR_INVALID, "jastrow.species out of range");
  }
  d.species = species.empty() ? nullptr : species.data();
  d.init_nuc = init_nuc.empty() ? nullptr : init_nuc.data();
  d.flags = flags;
  d.mo_rule = QMC_MO_FIXED;
  if (const JVal* mr = root.get("mo_rule")) {
    const std::string v = mr->k == JVal::STR ? mr->str : "";
    d.mo_rule = v == "fixed" ? QMC_MO_FIXED : v == "benzene_ch" ? QMC_MO_BENZENE_CH : -1;
    if (d.mo_rule < 0) return qmc_io_fail(QMC_ERR_INVALID, "mo_rule: expected \"fixed\" or \"benzene_ch\"");
  }
  return qmc_system_create(&d, precision, out);
}

}  // namespace

extern "C" int qmc_system_from_json(const char* text, int64_t len, int32_t precision, qmc_system** out) {
  if (!text || !out || len < 0) return qmc_io_fail(QMC_ERR_INVALID, "bad arguments");
  if (len == 0) len = (int64_t)std::strlen(text);
  if (len > (int64_t)64 << 20) return qmc_io_fail(QMC_ERR_INVALID, "description larger than 64 MiB");
  Reader r{text, text + len, {}};
  r.off0 = text;
  JVal root;
  if (!r.value(root, 0)) return qmc_io_fail(QMC_ERR_INVALID, ("JSON: " + r.err).c_str());
  r.ws();
  if (r.p != r.end) return qmc_io_fail(QMC_ERR_INVALID, "JSON: trailing characters");
  if (root.k != JVal::OBJ) return qmc_io_fail(QMC_ERR_INVALID, "JSON: the document must be an object");
  Conv cv;
  if (const JVal* s = root.get("system")) {
    if (s->k != JVal::STR) return qmc_io_fail(QMC_ERR_INVALID, "system: expected a name");
    return named(root, s->str, precision, out, cv);
  }
  if (root.get("orbital1") || root.get("H1")) return h2_layout(root, precision, out, cv);
  return generic(root, precision, out, cv);
}
