// Seeded synthetic RDF generators for the benchmark and test workloads.
//
// Id conventions follow Wukong's datagen (datagen/generate_data.cpp:357-373):
// 0 = __PREDICATE__, 1 = rdf:type, predicates/types from 2, normal vertices
// (entities and literals) from 2^17.
//
// University schema: per university 15-25 departments; per department 7-10
// full, 10-14 associate, 8-11 assistant professors and 5-7 lecturers,
// undergraduates 8-14x and graduate students 3-4x the faculty count, courses
// taught by the faculty, research groups and publications.  Every university
// is generated from its own RNG stream, so the output is identical for any
// thread count.
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "wk_common.hpp"
#include "wk_gpu.h"

namespace wk {
namespace {

struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed) : s(seed * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull) {}
  uint64_t next() {  // splitmix64
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  uint32_t range(uint32_t lo, uint32_t hi) {  // inclusive
    return lo + uint32_t(next() % uint64_t(hi - lo + 1));
  }
  double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
};

// ---- schema ---------------------------------------------------------------

enum : uint32_t {
  P_TYPE = TYPE_ID,
  P_MEMBER_OF = 2,
  P_SUB_ORG,
  P_ADVISOR,
  P_TAKES_COURSE,
  P_TEACHER_OF,
  P_WORKS_FOR,
  P_HEAD_OF,
  P_UG_DEGREE,
  P_MS_DEGREE,
  P_PHD_DEGREE,
  P_PUB_AUTHOR,
  P_TA_OF,
  P_RESEARCH_INTEREST,
  P_NAME,
  P_EMAIL,
  P_TELEPHONE,
  P_RESEARCH_PROJECT,  // 18 predicates with rdf:type
  T_UNIVERSITY,
  T_DEPARTMENT,
  T_FULL_PROF,
  T_ASSOC_PROF,
  T_ASSIST_PROF,
  T_LECTURER,
  T_UG_STUDENT,
  T_GRAD_STUDENT,
  T_RESEARCH_GROUP,
  T_COURSE,
  T_GRAD_COURSE,
  T_PUBLICATION,
  T_TA,
  T_RA,
  T_CHAIR,  // 15 types
  SCHEMA_END
};

struct NameId {
  const char* name;
  uint32_t id;
};
const NameId kSchema[] = {
    {"__PREDICATE__", PREDICATE_ID},
    {"rdf:type", P_TYPE},
    {"ub:memberOf", P_MEMBER_OF},
    {"ub:subOrganizationOf", P_SUB_ORG},
    {"ub:advisor", P_ADVISOR},
    {"ub:takesCourse", P_TAKES_COURSE},
    {"ub:teacherOf", P_TEACHER_OF},
    {"ub:worksFor", P_WORKS_FOR},
    {"ub:headOf", P_HEAD_OF},
    {"ub:undergraduateDegreeFrom", P_UG_DEGREE},
    {"ub:mastersDegreeFrom", P_MS_DEGREE},
    {"ub:doctoralDegreeFrom", P_PHD_DEGREE},
    {"ub:publicationAuthor", P_PUB_AUTHOR},
    {"ub:teachingAssistantOf", P_TA_OF},
    {"ub:researchInterest", P_RESEARCH_INTEREST},
    {"ub:name", P_NAME},
    {"ub:emailAddress", P_EMAIL},
    {"ub:telephone", P_TELEPHONE},
    {"ub:researchProject", P_RESEARCH_PROJECT},
    {"ub:University", T_UNIVERSITY},
    {"ub:Department", T_DEPARTMENT},
    {"ub:FullProfessor", T_FULL_PROF},
    {"ub:AssociateProfessor", T_ASSOC_PROF},
    {"ub:AssistantProfessor", T_ASSIST_PROF},
    {"ub:Lecturer", T_LECTURER},
    {"ub:UndergraduateStudent", T_UG_STUDENT},
    {"ub:GraduateStudent", T_GRAD_STUDENT},
    {"ub:ResearchGroup", T_RESEARCH_GROUP},
    {"ub:Course", T_COURSE},
    {"ub:GraduateCourse", T_GRAD_COURSE},
    {"ub:Publication", T_PUBLICATION},
    {"ub:TeachingAssistant", T_TA},
    {"ub:ResearchAssistant", T_RA},
    {"ub:Chair", T_CHAIR},
};

// Shared name literals ("FullProfessor3", "Course12", ...) repeat across
// departments, as LUBM's do; their ids form a fixed block at 2^17.
enum : uint32_t {
  L_FULL = 0, L_ASSOC, L_ASSIST, L_LECT, L_UG, L_GRAD, L_COURSE, L_GCOURSE, L_PUB, L_RESEARCH,
  L_DEPT, L_GROUP, L_PROJECT, L_KINDS
};
constexpr uint32_t kLitPerKind = 1024;
constexpr uint32_t kVidBase = uint32_t(1) << NBITS_IDX;
constexpr uint32_t kLocal = 0x80000000u;  // marks a university-local id before relocation

inline uint32_t name_lit(uint32_t kind, uint32_t i) { return kVidBase + kind * kLitPerKind + (i % kLitPerKind); }

struct UnivOut {
  std::vector<uint32_t> t;  // s,p,o with local ids tagged by kLocal
  uint32_t nlocal = 0;
};

void gen_university(uint32_t seed, uint32_t u, uint32_t n_univ, uint32_t univ_base, UnivOut& out) {
  Rng rng((uint64_t(seed) << 32) ^ (u + 1));
  auto& t = out.t;
  uint32_t next = 0;
  auto local = [&]() { return kLocal | next++; };
  auto add = [&](uint32_t s, uint32_t p, uint32_t o) {
    t.push_back(s);
    t.push_back(p);
    t.push_back(o);
  };
  auto rand_univ = [&]() { return univ_base + uint32_t(rng.next() % n_univ); };
  const uint32_t U = univ_base + u;
  add(U, P_TYPE, T_UNIVERSITY);
  add(U, P_NAME, local());  // unique literal "University<u>"
  const uint32_t ndept = rng.range(15, 25);
  for (uint32_t d = 0; d < ndept; d++) {
    const uint32_t D = local();
    add(D, P_TYPE, T_DEPARTMENT);
    add(D, P_SUB_ORG, U);
    add(D, P_NAME, name_lit(L_DEPT, d));
    const uint32_t nfull = rng.range(7, 10), nassoc = rng.range(10, 14), nassist = rng.range(8, 11),
                   nlect = rng.range(5, 7);
    const uint32_t nfac = nfull + nassoc + nassist + nlect;
    std::vector<uint32_t> profs, courses, gcourses;
    auto person = [&](uint32_t P, uint32_t kind, uint32_t i) {
      add(P, P_NAME, name_lit(kind, i));
      add(P, P_EMAIL, local());
      add(P, P_TELEPHONE, local());
    };
    uint32_t pub_no = 0;
    auto faculty = [&](uint32_t type, uint32_t kind, uint32_t n, uint32_t pub_lo, uint32_t pub_hi,
                       bool is_prof) {
      for (uint32_t i = 0; i < n; i++) {
        const uint32_t F = local();
        add(F, P_TYPE, type);
        add(F, P_WORKS_FOR, D);
        person(F, kind, i);
        add(F, P_RESEARCH_INTEREST, name_lit(L_RESEARCH, rng.range(0, 29)));
        add(F, P_UG_DEGREE, rand_univ());
        add(F, P_MS_DEGREE, rand_univ());
        add(F, P_PHD_DEGREE, rand_univ());
        const uint32_t nc = rng.range(1, 2), ngc = rng.range(1, 2);
        for (uint32_t c = 0; c < nc; c++) {
          const uint32_t C = local();
          add(C, P_TYPE, T_COURSE);
          add(C, P_NAME, name_lit(L_COURSE, uint32_t(courses.size())));
          add(F, P_TEACHER_OF, C);
          courses.push_back(C);
        }
        for (uint32_t c = 0; c < ngc; c++) {
          const uint32_t C = local();
          add(C, P_TYPE, T_GRAD_COURSE);
          add(C, P_NAME, name_lit(L_GCOURSE, uint32_t(gcourses.size())));
          add(F, P_TEACHER_OF, C);
          gcourses.push_back(C);
        }
        const uint32_t npub = rng.range(pub_lo, pub_hi);
        for (uint32_t k = 0; k < npub; k++) {
          const uint32_t Pb = local();
          add(Pb, P_TYPE, T_PUBLICATION);
          add(Pb, P_NAME, name_lit(L_PUB, pub_no++));
          add(Pb, P_PUB_AUTHOR, F);
        }
        if (type == T_FULL_PROF && i == 0) {
          add(F, P_TYPE, T_CHAIR);
          add(F, P_HEAD_OF, D);
        }
        if (is_prof) profs.push_back(F);
      }
    };
    faculty(T_FULL_PROF, L_FULL, nfull, 15, 20, true);
    faculty(T_ASSOC_PROF, L_ASSOC, nassoc, 10, 18, true);
    faculty(T_ASSIST_PROF, L_ASSIST, nassist, 5, 10, true);
    faculty(T_LECTURER, L_LECT, nlect, 0, 5, false);
    const uint32_t nug = nfac * rng.range(8, 14);
    for (uint32_t i = 0; i < nug; i++) {
      const uint32_t S = local();
      add(S, P_TYPE, T_UG_STUDENT);
      add(S, P_MEMBER_OF, D);
      person(S, L_UG, i);
      const uint32_t nt = rng.range(2, 4);
      for (uint32_t k = 0; k < nt; k++) add(S, P_TAKES_COURSE, courses[rng.next() % courses.size()]);
      if (rng.range(0, 4) == 0) add(S, P_ADVISOR, profs[rng.next() % profs.size()]);
    }
    const uint32_t ngr = nfac * rng.range(3, 4);
    std::vector<uint32_t> grads;
    for (uint32_t i = 0; i < ngr; i++) {
      const uint32_t S = local();
      grads.push_back(S);
      add(S, P_TYPE, T_GRAD_STUDENT);
      add(S, P_MEMBER_OF, D);
      person(S, L_GRAD, i);
      add(S, P_UG_DEGREE, rand_univ());
      add(S, P_ADVISOR, profs[rng.next() % profs.size()]);
      const uint32_t nt = rng.range(1, 3);
      for (uint32_t k = 0; k < nt; k++) add(S, P_TAKES_COURSE, gcourses[rng.next() % gcourses.size()]);
      if (rng.range(0, 4) == 0) {
        add(S, P_TYPE, T_TA);
        add(S, P_TA_OF, courses[rng.next() % courses.size()]);
      }
      if (rng.range(0, 3) == 0) add(S, P_TYPE, T_RA);
    }
    const uint32_t ngroups = rng.range(10, 20);
    for (uint32_t g = 0; g < ngroups; g++) {
      const uint32_t G = local();
      add(G, P_TYPE, T_RESEARCH_GROUP);
      add(G, P_SUB_ORG, D);
      add(G, P_RESEARCH_PROJECT, name_lit(L_PROJECT, g));
    }
    // graduate students co-author some of the department's publications
    // (authorship added to fresh publication vertices keeps ids contiguous)
    const uint32_t ncoauth = ngr / 2;
    for (uint32_t k = 0; k < ncoauth; k++) {
      const uint32_t Pb = local();
      add(Pb, P_TYPE, T_PUBLICATION);
      add(Pb, P_NAME, name_lit(L_PUB, pub_no++));
      add(Pb, P_PUB_AUTHOR, grads[rng.next() % grads.size()]);
      add(Pb, P_PUB_AUTHOR, profs[rng.next() % profs.size()]);
    }
  }
  out.nlocal = next;
}

}  // namespace
}  // namespace wk

using namespace wk;

extern "C" uint32_t wk_univ_id(const char* name) {
  if (!name) return 0;
  for (const auto& e : kSchema)
    if (std::strcmp(e.name, name) == 0) return e.id;
  return 0;
}

extern "C" int wk_gen_university(uint32_t seed, uint32_t n_univ, uint32_t** triples, uint64_t* n) {
  if (!triples || !n || n_univ == 0) return WK_E_INVALID_ARG;
  *triples = nullptr;
  *n = 0;
  const uint32_t univ_base = kVidBase + L_KINDS * kLitPerKind;
  const uint32_t local_base = univ_base + n_univ;
  std::vector<UnivOut> outs(n_univ);
#pragma omp parallel for schedule(dynamic, 1)
  for (uint32_t u = 0; u < n_univ; u++) gen_university(seed, u, n_univ, univ_base, outs[u]);
  std::vector<uint64_t> toff(n_univ + 1, 0);
  std::vector<uint64_t> idoff(n_univ + 1, 0);
  for (uint32_t u = 0; u < n_univ; u++) {
    toff[u + 1] = toff[u] + outs[u].t.size();
    idoff[u + 1] = idoff[u] + outs[u].nlocal;
  }
  if (uint64_t(local_base) + idoff[n_univ] >= (uint64_t(1) << 31)) return WK_E_OVERFLOW;
  uint32_t* buf = static_cast<uint32_t*>(std::malloc(std::max<uint64_t>(toff[n_univ], 1) * sizeof(uint32_t)));
  if (!buf) return WK_E_NOMEM_HOST;
#pragma omp parallel for schedule(dynamic, 1)
  for (uint32_t u = 0; u < n_univ; u++) {
    const uint32_t base = uint32_t(local_base + idoff[u]);
    uint32_t* dst = buf + toff[u];
    const auto& src = outs[u].t;
    for (size_t i = 0; i < src.size(); i++) {
      const uint32_t v = src[i];
      dst[i] = (v & kLocal) ? base + (v & ~kLocal) : v;
    }
    std::vector<uint32_t>().swap(outs[u].t);
  }
  *triples = buf;
  *n = toff[n_univ] / 3;
  return WK_OK;
}

// Power-law graph: vertex v (id 2^17 + v) gets one uniform type and an
// out-degree drawn from Pareto(alpha, 1) truncated at max_degree and scaled so
// the edge total matches; each edge's predicate follows Zipf(zipf_s) over
// n_preds predicates and its object is uniform.  Vertices are processed in
// fixed chunks with per-chunk RNG streams (thread-count independent).
extern "C" int wk_gen_powerlaw(uint32_t seed, uint64_t n_vertices, uint64_t n_triples, uint32_t n_preds,
                               uint32_t n_types, double zipf_s, double pareto_alpha, uint64_t max_degree,
                               uint32_t** triples, uint64_t* n) {
  if (!triples || !n || n_vertices == 0 || n_preds == 0 || n_types == 0 || pareto_alpha <= 0 ||
      max_degree == 0)
    return WK_E_INVALID_ARG;
  *triples = nullptr;
  *n = 0;
  if (2ull + n_preds + n_types >= (1ull << NBITS_IDX)) return WK_E_INVALID_ARG;
  if (uint64_t(kVidBase) + n_vertices >= (1ull << 32)) return WK_E_OVERFLOW;
  const uint64_t n_edges = n_triples > n_vertices ? n_triples - n_vertices : 0;
  const uint32_t pred_base = 2, type_base = 2 + n_preds;
  // Zipf CDF over predicates
  std::vector<double> cdf(n_preds);
  double acc = 0;
  for (uint32_t i = 0; i < n_preds; i++) {
    acc += 1.0 / std::pow(double(i + 1), zipf_s);
    cdf[i] = acc;
  }
  for (auto& c : cdf) c /= acc;
  constexpr uint64_t CH = 1 << 16;
  const uint64_t nch = (n_vertices + CH - 1) / CH;
  // pass 1: raw Pareto degrees
  std::vector<uint32_t> deg(n_vertices);
  std::vector<double> chsum(nch, 0.0);
#pragma omp parallel for schedule(dynamic, 4)
  for (uint64_t c = 0; c < nch; c++) {
    Rng rng((uint64_t(seed) << 32) ^ (c * 2 + 1));
    double s = 0;
    for (uint64_t v = c * CH; v < std::min(n_vertices, (c + 1) * CH); v++) {
      double u = 1.0 - rng.uniform();
      double d = std::pow(u, -1.0 / pareto_alpha);
      if (d > double(max_degree)) d = double(max_degree);
      deg[v] = uint32_t(d);
      s += deg[v];
    }
    chsum[c] = s;
  }
  double total = 0;
  for (double s : chsum) total += s;
  const double scale = total > 0 ? double(n_edges) / total : 0.0;
  std::vector<uint64_t> voff(nch + 1, 0);
#pragma omp parallel for schedule(dynamic, 4)
  for (uint64_t c = 0; c < nch; c++) {
    Rng rng((uint64_t(seed) << 32) ^ (c * 2 + 2));
    uint64_t s = 0;
    for (uint64_t v = c * CH; v < std::min(n_vertices, (c + 1) * CH); v++) {
      const double x = deg[v] * scale;
      uint64_t d = uint64_t(x);
      if (rng.uniform() < x - double(d)) d++;
      if (d > max_degree) d = max_degree;
      deg[v] = uint32_t(d);
      s += d + 1;  // + type triple
    }
    voff[c + 1] = s;
  }
  for (uint64_t c = 0; c < nch; c++) voff[c + 1] += voff[c];
  const uint64_t nt = voff[nch];
  uint32_t* buf = static_cast<uint32_t*>(std::malloc(std::max<uint64_t>(nt, 1) * 3 * sizeof(uint32_t)));
  if (!buf) return WK_E_NOMEM_HOST;
#pragma omp parallel for schedule(dynamic, 4)
  for (uint64_t c = 0; c < nch; c++) {
    Rng rng((uint64_t(seed) << 32) ^ (c * 2 + 3) ^ 0xABCDEF0000000000ull);
    uint32_t* dst = buf + 3 * voff[c];
    for (uint64_t v = c * CH; v < std::min(n_vertices, (c + 1) * CH); v++) {
      const uint32_t s = uint32_t(kVidBase + v);
      *dst++ = s;
      *dst++ = TYPE_ID;
      *dst++ = type_base + uint32_t(rng.next() % n_types);
      for (uint32_t k = 0; k < deg[v]; k++) {
        const double r = rng.uniform();
        const uint32_t pi = uint32_t(std::lower_bound(cdf.begin(), cdf.end(), r) - cdf.begin());
        *dst++ = s;
        *dst++ = pred_base + std::min(pi, n_preds - 1);
        *dst++ = uint32_t(kVidBase + rng.next() % n_vertices);
      }
    }
  }
  *triples = buf;
  *n = nt;
  return WK_OK;
}

extern "C" void wk_free(void* p) { std::free(p); }
