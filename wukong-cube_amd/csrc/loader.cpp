// wk_store_load_dir: the normal-triple part of BaseLoader::load
// (loader/base_loader.hpp:536-576) over a POSIX directory
// (PosixLoader::list_files, loader/posix_loader.hpp:49-69).
//
// - Files: every entry of the directory whose name starts with "id_"; names
//   starting with '.' are skipped, and the list is sorted.
// - Format: whitespace-separated unsigned integers read three at a time, as
//   `while (file >> s >> p >> o)` does (base_loader.hpp:245).  Reading a
//   file stops at the first token that is not an unsigned 32-bit number, and
//   an incomplete last triple is dropped.
// - Share of server `sid` among `n`: out-edges are the triples with
//   s % n == sid, in-edges those with o % n == sid (PARTITION,
//   core/store/vertex.hpp:59; aggregate_data, base_loader.hpp:342-397).
// - Sorting, dedup and insertion are wk_store_build's (sort_normal_triples,
//   base_loader.hpp:481-496; SegmentRDFGraph::init_gstore).
//
// Files are read whole and parsed by hand, one file per thread.
#include <dirent.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "store.hpp"
#include "wk_gpu.h"

namespace {

bool read_file(const std::string& path, std::string* buf) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) return false;
  if (std::fseek(f, 0, SEEK_END) == 0) {
    const long size = std::ftell(f);
    if (size > 0) buf->reserve(size_t(size));
    std::rewind(f);
  }
  char chunk[1 << 16];
  size_t n;
  while ((n = std::fread(chunk, 1, sizeof(chunk), f)) > 0) buf->append(chunk, n);
  const bool ok = !std::ferror(f);
  std::fclose(f);
  return ok;
}

// Whitespace and digits of the "C" locale, which the reference's streams use.
inline bool is_space(char c) { return c == ' ' || (c >= '\t' && c <= '\r'); }
inline bool is_digit(char c) { return unsigned(c - '0') < 10u; }

// Unsigned 32-bit numbers of `text` up to the first read that fails: after
// optional whitespace, an optional '+' and one or more digits whose value
// fits 32 bits.  A '-' sign fails the read (ids are non-negative).
void parse_ids(const std::string& text, std::vector<uint32_t>* out) {
  const char* p = text.data();
  const char* e = p + text.size();
  out->reserve(text.size() / 8);
  for (;;) {
    while (p < e && is_space(*p)) ++p;
    if (p == e) return;
    if (*p == '+') ++p;
    if (p == e || !is_digit(*p)) return;
    uint64_t v = 0;
    while (p < e && is_digit(*p)) {
      v = v * 10 + uint64_t(*p - '0');
      if (v > UINT32_MAX) return;
      ++p;
    }
    // as with operator>>, a number ends at the first non-digit: "12abc" reads
    // 12 and then fails on "abc"; "12+3" reads 12 and 3
    out->push_back(uint32_t(v));
  }
}

}  // namespace

extern "C" int wk_store_load_dir(const char* dir, uint32_t server, uint32_t nservers, const wk_store_opts* opts,
                                 wk_store** out, uint64_t* triples_read) {
  if (!dir || !out || nservers == 0 || server >= nservers) return WK_E_INVALID_ARG;
  *out = nullptr;
  if (triples_read) *triples_read = 0;
  std::string base(dir);
  if (base.empty() || base.back() != '/') base += '/';
  DIR* d = opendir(base.c_str());
  if (!d) return WK_E_FILE_NOT_FOUND;
  std::vector<std::string> names;
  while (struct dirent* ent = readdir(d)) {
    if (ent->d_name[0] == '.') continue;
    if (std::strncmp(ent->d_name, "id_", 3) == 0) names.push_back(base + ent->d_name);
  }
  closedir(d);
  std::sort(names.begin(), names.end());

  std::vector<std::vector<uint32_t>> ids(names.size());
  int failed = 0, nomem = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : failed, nomem)
  for (size_t i = 0; i < names.size(); i++) {
    try {
      std::string text;
      if (!read_file(names[i], &text)) {
        failed++;
        continue;
      }
      parse_ids(text, &ids[i]);
      ids[i].resize(ids[i].size() / 3 * 3);  // an incomplete last triple is not read
    } catch (const std::bad_alloc&) {
      nomem++;
    }
  }
  if (nomem) return WK_E_NOMEM_HOST;
  if (failed) return WK_E_FILE_NOT_FOUND;

  // this server's out-edges (s % n) and in-edges (o % n), file order kept
  uint64_t total = 0;
  for (auto& v : ids) total += v.size() / 3;
  std::vector<uint32_t> pso, pos;
  try {
    if (nservers == 1) {
      pso.reserve(total * 3);
      for (auto& v : ids) {
        pso.insert(pso.end(), v.begin(), v.end());
        std::vector<uint32_t>().swap(v);
      }
    } else {
      for (auto& v : ids) {
        for (size_t k = 0; k < v.size(); k += 3) {
          if (v[k] % nservers == server) pso.insert(pso.end(), &v[k], &v[k] + 3);
          if (v[k + 2] % nservers == server) pos.insert(pos.end(), &v[k], &v[k] + 3);
        }
        std::vector<uint32_t>().swap(v);
      }
    }
  } catch (const std::bad_alloc&) {
    return WK_E_NOMEM_HOST;
  }
  if (triples_read) *triples_read = total;
  const std::vector<uint32_t>& in_edges = nservers == 1 ? pso : pos;
  return wk_store_build(pso.data(), pso.size() / 3, in_edges.data(), in_edges.size() / 3, opts, out);
}
