// The reference's wire form of SPARQLQuery (a Bundle string holding a boost
// binary archive) and a GPU agent that takes a request bundle and returns
// the reply bundle.
//
// Byte layout, as boost 1.67's binary_oarchive writes it on x86-64 for the
// save() functions of core/sparql/query.hpp:1086-1276 (every class
// object_serializable, so no class preamble; TRDF_MODE off):
//   header   u64 22, "serialization::archive", u16 library version,
//            u8 sizeof(int) = 4, sizeof(long) = 8, sizeof(float) = 4,
//            sizeof(double) = 8, i32 1 (byte order)
//   int, enum, ssid_t, unsigned   4 bytes;  bool, char   1 byte
//   vector<int32/uint32>          u64 count, then the elements as one array
//   vector<bool>                  u64 count, one byte per element
//   vector<class>, set<int32>     u64 count, u32 item version (0), elements
//   an optional member            i8 flag: 0 = present (`occupied`), 1 = absent
//   Result::gpu (USE_GPU)         GPUResult has the default class level, so
//            its first (only) instance in an archive is preceded by its class
//            info: u8 tracking (0) and u32 class version (0); then
//            u64 result_buf_nelems, i32 col_num
#include <algorithm>
#include <cstring>
#include <new>
#include <type_traits>
#include <vector>

#include "device.hpp"

using namespace wk;

namespace {

constexpr uint32_t kReqSparql = 0;  // req_type SPARQL_QUERY (core/common/type.hpp:52-55)
constexpr uint16_t kDefaultLibraryVersion = 16;
constexpr char kSignature[] = "serialization::archive";
constexpr int8_t kOccupied = 0, kEmpty = 1;  // query.hpp:1086-1087
constexpr int32_t kNoResult = 0xffff;       // NO_RESULT, NBITS_COL = 16 (query.hpp:60-63)

struct Out {
  std::vector<uint8_t> b;
  template <class T>
  void put(T v) {
    static_assert(std::is_trivially_copyable<T>::value, "plain value");
    const uint8_t* p = reinterpret_cast<const uint8_t*>(&v);
    b.insert(b.end(), p, p + sizeof(T));
  }
  void bytes(const void* p, uint64_t n) {
    const uint8_t* q = static_cast<const uint8_t*>(p);
    if (n) b.insert(b.end(), q, q + n);
  }
  void flag(bool present) { put<int8_t>(present ? kOccupied : kEmpty); }
};

struct In {
  const uint8_t* p;
  uint64_t n, pos = 0;
  bool ok = true;
  template <class T>
  T get() {
    T v{};
    if (!ok || n - pos < sizeof(T)) {
      ok = false;
      return v;
    }
    std::memcpy(&v, p + pos, sizeof(T));
    pos += sizeof(T);
    return v;
  }
  // a collection count whose elements of `elem` bytes each must still fit
  uint64_t count(uint64_t elem) {
    const uint64_t c = get<uint64_t>();
    if (!ok || (elem && c > (n - pos) / elem)) ok = false;
    return ok ? c : 0;
  }
  template <class T>
  void array(std::vector<T>& v) {
    const uint64_t c = count(sizeof(T));
    if (!ok) return;
    v.resize(c);
    if (c) std::memcpy(v.data(), p + pos, c * sizeof(T));
    pos += c * sizeof(T);
  }
  bool present() { return get<int8_t>() == kOccupied; }  // any other byte reads as absent, as load() does
};

// A decoded message: the C view first, then the storage its pointers use.
struct Holder {
  wk_sparql_msg m;
  std::vector<wk_wire_pattern> patterns;
  std::vector<int32_t> new_vars, order_vars, order_desc, required, v2c;
  std::vector<uint8_t> opt_matched;
  std::vector<uint32_t> table;

  void point() {
    m.patterns = patterns.data();
    m.npatterns = uint32_t(patterns.size());
    m.new_vars = new_vars.data();
    m.nnew_vars = uint32_t(new_vars.size());
    m.order_vars = order_vars.data();
    m.order_desc = order_desc.data();
    m.norders = uint32_t(order_vars.size());
    m.required_vars = required.data();
    m.nrequired = uint32_t(required.size());
    m.v2c_map = v2c.data();
    m.nv2c = uint32_t(v2c.size());
    m.optional_matched = opt_matched.data();
    m.noptional_matched = opt_matched.size();
    m.table = table.data();
    m.ntable = table.size();
  }
};
static_assert(std::is_standard_layout<Holder>::value, "the C view must sit at the holder's address");

uint16_t library_version(const wk_wire_opts* o) {
  return (o && o->library_version) ? uint16_t(o->library_version) : kDefaultLibraryVersion;
}

int check_msg(const wk_sparql_msg& m) {
  if ((m.npatterns && !m.patterns) || (m.nnew_vars && !m.new_vars) ||
      (m.norders && (!m.order_vars || !m.order_desc)) || (m.nrequired && !m.required_vars) ||
      (m.nv2c && !m.v2c_map) || (m.noptional_matched && !m.optional_matched) || (m.ntable && !m.table))
    return WK_E_INVALID_ARG;
  return WK_OK;
}

// save(SPARQLQuery) and the saves it calls (query.hpp:1089-1100, 1113-1132,
// 1155-1181, 1208-1243), preceded by the Bundle's type word.
void encode(const wk_sparql_msg& m, uint16_t lv, bool use_gpu, Out& o) {
  o.put<uint32_t>(kReqSparql);
  o.put<uint64_t>(sizeof(kSignature) - 1);
  o.bytes(kSignature, sizeof(kSignature) - 1);
  o.put<uint16_t>(lv);
  o.put<uint8_t>(4);  // sizeof(int)
  o.put<uint8_t>(8);  // sizeof(long)
  o.put<uint8_t>(4);  // sizeof(float)
  o.put<uint8_t>(8);  // sizeof(double)
  o.put<int32_t>(1);  // byte-order probe

  const int32_t head1[] = {m.q_type,   m.qid,    m.pqid,         m.pg_type,   m.state,    m.dev_type,
                           m.job_type, m.priority, m.mt_factor, m.mt_tid, m.pattern_step, m.local_var};
  for (int32_t v : head1) o.put<int32_t>(v);
  o.put<uint8_t>(m.corun_enabled != 0);
  o.put<int32_t>(m.corun_step);
  o.put<int32_t>(m.fetch_step);
  o.put<uint8_t>(m.union_done != 0);
  o.put<int32_t>(m.optional_step);
  o.put<int32_t>(m.limit);
  o.put<uint32_t>(m.offset);
  o.put<uint8_t>(m.distinct != 0);

  // PatternGroup
  o.put<uint64_t>(m.npatterns);
  o.put<uint32_t>(0);  // item version
  for (uint32_t i = 0; i < m.npatterns; i++) {
    const wk_wire_pattern& p = m.patterns[i];
    o.put<int32_t>(p.subject);
    o.put<int32_t>(p.predicate);
    o.put<int32_t>(p.object);
    o.put<int32_t>(p.direction);
    o.put<int8_t>(int8_t(p.pred_type));
  }
  std::vector<int32_t> nv(m.new_vars, m.new_vars + m.nnew_vars);  // a std::set: ascending, unique
  std::sort(nv.begin(), nv.end());
  nv.erase(std::unique(nv.begin(), nv.end()), nv.end());
  o.put<uint64_t>(nv.size());
  o.put<uint32_t>(0);
  for (int32_t v : nv) o.put<int32_t>(v);
  o.flag(false);  // filters
  o.flag(false);  // optional
  o.flag(false);  // unions

  // orders
  o.flag(m.norders > 0);
  if (m.norders) {
    o.put<uint64_t>(m.norders);
    o.put<uint32_t>(0);
    for (uint32_t i = 0; i < m.norders; i++) {
      o.put<int32_t>(m.order_vars[i]);
      o.put<uint8_t>(m.order_desc[i] != 0);
    }
  }

  // Result
  o.put<int32_t>(m.col_num);
  o.put<int32_t>(m.row_num);
  o.put<int32_t>(m.attr_col_num);
  o.put<int32_t>(m.status_code);
  o.put<uint8_t>(m.blind != 0);
  o.put<int32_t>(m.nvars);
  o.put<uint64_t>(m.nrequired);
  o.bytes(m.required_vars, uint64_t(m.nrequired) * 4);
  o.put<uint64_t>(m.nv2c);
  o.bytes(m.v2c_map, uint64_t(m.nv2c) * 4);
  o.put<uint64_t>(m.noptional_matched);
  for (uint64_t i = 0; i < m.noptional_matched; i++) o.put<uint8_t>(m.optional_matched[i] != 0);
  o.flag(m.row_num > 0);
  if (m.row_num > 0) {
    o.put<uint64_t>(m.ntable);
    o.bytes(m.table, m.ntable * 4);
    o.put<uint64_t>(0);  // attr_res_table: no attribute results on this path
    o.put<uint32_t>(0);
  }
  if (use_gpu) {
    o.put<uint8_t>(0);   // tracking: GPUResult is never saved through a pointer
    o.put<uint32_t>(0);  // class version
    o.put<uint64_t>(m.gpu_nelems);
    o.put<int32_t>(m.gpu_col_num);
  }
}

// load(SPARQLQuery) and the loads it calls (query.hpp:1101-1111, 1133-1153,
// 1182-1206, 1244-1276).
int decode(const uint8_t* src, uint64_t len, bool use_gpu, Holder& h) {
  In in{src, len};
  wk_sparql_msg& m = h.m;
  std::memset(&m, 0, sizeof(m));
  if (in.get<uint32_t>() != kReqSparql || !in.ok) return WK_E_WIRE;
  const uint64_t sl = in.get<uint64_t>();
  if (!in.ok || sl != sizeof(kSignature) - 1 || len - in.pos < sl ||
      std::memcmp(src + in.pos, kSignature, sl) != 0)
    return WK_E_WIRE;
  in.pos += sl;
  const uint16_t lv = in.get<uint16_t>();
  // collection counts are 8 bytes and class versions 4 from library version 7 on
  if (!in.ok || lv < 7) return WK_E_WIRE;
  const uint8_t si = in.get<uint8_t>(), sl2 = in.get<uint8_t>(), sf = in.get<uint8_t>(), sd = in.get<uint8_t>();
  if (!in.ok || si != 4 || sl2 != 8 || sf != 4 || sd != 8 || in.get<int32_t>() != 1) return WK_E_WIRE;

  int32_t* head1[] = {&m.q_type,   &m.qid,      &m.pqid,      &m.pg_type, &m.state,        &m.dev_type,
                      &m.job_type, &m.priority, &m.mt_factor, &m.mt_tid,  &m.pattern_step, &m.local_var};
  for (int32_t* v : head1) *v = in.get<int32_t>();
  m.corun_enabled = in.get<uint8_t>() != 0;
  m.corun_step = in.get<int32_t>();
  m.fetch_step = in.get<int32_t>();
  m.union_done = in.get<uint8_t>() != 0;
  m.optional_step = in.get<int32_t>();
  m.limit = in.get<int32_t>();
  m.offset = in.get<uint32_t>();
  m.distinct = in.get<uint8_t>() != 0;

  const uint64_t np = in.count(17);
  (void)in.get<uint32_t>();  // item version
  if (!in.ok) return WK_E_WIRE;
  h.patterns.resize(np);
  for (auto& p : h.patterns) {
    p.subject = in.get<int32_t>();
    p.predicate = in.get<int32_t>();
    p.object = in.get<int32_t>();
    p.direction = in.get<int32_t>();
    p.pred_type = in.get<int8_t>();
  }
  const uint64_t nn = in.count(4);
  (void)in.get<uint32_t>();
  if (!in.ok) return WK_E_WIRE;
  h.new_vars.resize(nn);
  for (auto& v : h.new_vars) v = in.get<int32_t>();
  const bool filters = in.present();
  if (!in.ok) return WK_E_WIRE;
  if (filters) return WK_E_UNKNOWN_FILTER;
  const bool optional = in.present();
  if (!in.ok) return WK_E_WIRE;
  if (optional) return WK_E_UNSUPPORT_UNION;
  const bool unions = in.present();
  if (!in.ok) return WK_E_WIRE;
  if (unions) return WK_E_UNSUPPORT_UNION;

  if (in.present()) {
    const uint64_t no = in.count(5);
    (void)in.get<uint32_t>();
    if (!in.ok) return WK_E_WIRE;
    h.order_vars.resize(no);
    h.order_desc.resize(no);
    for (uint64_t i = 0; i < no; i++) {
      h.order_vars[i] = in.get<int32_t>();
      h.order_desc[i] = in.get<uint8_t>() != 0;
    }
  }

  m.col_num = in.get<int32_t>();
  m.row_num = in.get<int32_t>();
  m.attr_col_num = in.get<int32_t>();
  m.status_code = in.get<int32_t>();
  m.blind = in.get<uint8_t>() != 0;
  m.nvars = in.get<int32_t>();
  in.array(h.required);
  in.array(h.v2c);
  in.array(h.opt_matched);
  for (auto& b : h.opt_matched) b = b != 0;
  if (!in.ok) return WK_E_WIRE;
  if (in.present()) {
    in.array(h.table);
    const uint64_t na = in.count(5);  // a variant is at least its 4-byte `which` and a value
    (void)in.get<uint32_t>();
    if (!in.ok) return WK_E_WIRE;
    if (na) return WK_E_ATTR_DISABLE;
  }
  if (use_gpu) {
    (void)in.get<uint8_t>();   // tracking
    (void)in.get<uint32_t>();  // class version
    m.gpu_nelems = in.get<uint64_t>();
    m.gpu_col_num = in.get<int32_t>();
  }
  if (!in.ok || h.patterns.size() > UINT32_MAX || h.table.size() > UINT64_MAX / 4) return WK_E_WIRE;
  h.point();
  return WK_OK;
}

// The device history table of a request, run from its pattern_step to the
// end through the per-pattern calls (GPUEngine::execute_one_pattern,
// gpu/gpu_engine.hpp:265-338).  `t` holds the request's table on entry.
int continue_patterns(wk_dev* dev, const wk_sparql_msg& m, std::vector<int32_t>& v2c, wk_table** t) {
  for (uint32_t k = uint32_t(m.pattern_step); k < m.npatterns; k++) {
    const wk_wire_pattern& p = m.patterns[k];
    if (p.predicate < 0 || (p.direction != WK_IN && p.direction != WK_OUT)) return WK_E_UNKNOWN_PATTERN;
    auto idx = [&](int32_t x) { return -(int64_t(x) + 1); };
    for (int32_t x : {p.subject, p.object})
      if (x < 0 && idx(x) >= m.nvars) return WK_E_VERTEX_INVALID;
    wk_table* next = nullptr;
    int rc;
    if (p.subject >= 0) {
      // a constant start over the carried table: the first pattern from an
      // index (start_from_index, query.hpp:816-838) is index_to_known when
      // its object is bound (sparql.hpp:83-137, dispatched at :945-950);
      // otherwise const_to_known (:139-188).  An unbound object would start
      // a new table, which must be the first pattern.
      const bool index_start = k == 0 && is_tpid(p.subject);
      if (index_start && p.predicate != int32_t(PREDICATE_ID) && p.predicate != int32_t(TYPE_ID))
        return WK_E_OBJ_ERROR;
      if (p.object >= 0) return index_start ? WK_E_VERTEX_INVALID : WK_E_UNKNOWN_PATTERN;
      const int32_t ec = v2c[size_t(idx(p.object))];
      if (ec < 0) return WK_E_FIRST_PATTERN;
      rc = index_start ? wk_expand_c2k(dev, *t, 0, uint32_t(p.subject), p.direction, uint32_t(ec), &next)
                       : wk_expand_c2k(dev, *t, uint32_t(p.subject), uint32_t(p.predicate), p.direction,
                                       uint32_t(ec), &next);
      if (rc) return rc;
      (void)wk_table_free(dev, *t);
      *t = next;
      continue;
    }
    // an unbound start cannot follow a history table
    if (v2c[size_t(idx(p.subject))] < 0) return WK_E_FIRST_PATTERN;
    const uint32_t sc = uint32_t(v2c[size_t(idx(p.subject))]);
    if (p.object >= 0) {
      rc = wk_expand_k2c(dev, *t, sc, uint32_t(p.predicate), p.direction, uint32_t(p.object), &next);
    } else if (v2c[size_t(idx(p.object))] >= 0) {
      rc = wk_expand_k2k(dev, *t, sc, uint32_t(p.predicate), p.direction, uint32_t(v2c[size_t(idx(p.object))]),
                         &next);
    } else {
      rc = wk_expand_k2u(dev, *t, sc, uint32_t(p.predicate), p.direction, &next);
      if (rc == WK_OK) v2c[size_t(idx(p.object))] = int32_t((*t)->cols);
    }
    if (rc) return rc;
    (void)wk_table_free(dev, *t);
    *t = next;
  }
  return WK_OK;
}

// failures of the device or the host, not of the query: the call returns them
bool is_device_error(int rc) {
  return rc == WK_E_HIP || rc == WK_E_NO_DEVICE || rc == WK_E_NOMEM_HOST || rc == WK_E_NOMEM_BUDGET;
}

}  // namespace

extern "C" {

int wk_wire_encode(const wk_sparql_msg* m, const wk_wire_opts* o, uint8_t* dst, uint64_t cap, uint64_t* len) {
  if (!m || !len) return WK_E_INVALID_ARG;
  if (int rc = check_msg(*m)) return rc;
  Out out;
  encode(*m, library_version(o), o && o->use_gpu, out);
  *len = out.b.size();
  if (!dst || cap < out.b.size()) return WK_E_INVALID_ARG;
  std::memcpy(dst, out.b.data(), out.b.size());
  return WK_OK;
}

int wk_wire_decode(const uint8_t* src, uint64_t len, const wk_wire_opts* o, wk_sparql_msg** out) {
  if (!out || (len && !src)) return WK_E_INVALID_ARG;
  *out = nullptr;
  Holder* h = new (std::nothrow) Holder();
  if (!h) return WK_E_NOMEM_HOST;
  const int rc = decode(src, len, o && o->use_gpu, *h);
  if (rc) {
    delete h;
    return rc;
  }
  *out = &h->m;
  return WK_OK;
}

void wk_wire_free(wk_sparql_msg* m) { delete reinterpret_cast<Holder*>(m); }

int wk_agent_execute(wk_dev* dev, const uint8_t* req, uint64_t len, const wk_wire_opts* o, uint64_t* reply_len) {
  if (!dev || !reply_len || (len && !req)) return WK_E_INVALID_ARG;
  if (int drc = drain_batches(dev)) return drc;
  *reply_len = 0;
  dev->agent_reply.clear();
  Holder q;
  if (decode(req, len, o && o->use_gpu, q) != WK_OK) return WK_E_WIRE;
  wk_sparql_msg& m = q.m;
  if (m.qid == -1) m.qid = ++dev->agent_qid;  // Coder::get_and_inc_qid (gpu_agent.hpp:174)

  Holder r = q;  // the reply starts as the request
  r.point();
  int status = WK_OK;
  if (m.dev_type != WK_DEV_GPU || m.state == WK_SQ_REPLY || m.npatterns == 0 || m.pattern_step < 0 ||
      uint32_t(m.pattern_step) > m.npatterns)
    status = WK_E_SETTING;
  if (!status && m.nvars > WK_MAX_VARS) status = WK_E_TOO_MANY_VARS;
  if (!status && m.nvars < 0) status = WK_E_VERTEX_INVALID;

  std::vector<int32_t> v2c(size_t(std::max(m.nvars, 0)), -1);
  wk_table* t = nullptr;
  wk_table view{};  // the handle's last result, borrowed
  const wk_table* result = nullptr;
  int rc = WK_OK;
  if (!status && m.pattern_step == 0 && m.col_num == 0) {
    // the whole pattern list in one device pass
    std::vector<wk_pattern> pats(m.npatterns);
    for (uint32_t i = 0; i < m.npatterns; i++)
      pats[i] = wk_pattern{m.patterns[i].subject, m.patterns[i].predicate, m.patterns[i].object,
                           m.patterns[i].direction};
    const wk_query wq{pats.data(), m.npatterns, m.nvars};
    uint64_t rows = 0;
    uint32_t cols = 0;
    rc = wk_query_run(dev, &wq, &rows, &cols, v2c.data());
    if (rc == WK_OK) {
      view.d = const_cast<uint32_t*>(dev->last_ptr);
      view.rows = rows;
      view.cols = cols;
      result = &view;
    }
  } else if (!status) {
    // GPUEngine::load_result_buf: continue from the request's history table
    if (m.col_num <= 0 || m.ntable % uint64_t(m.col_num)) {
      status = WK_E_SETTING;
    } else {
      for (uint32_t i = 0; i < m.nv2c && i < v2c.size(); i++) {
        const int32_t c = m.v2c_map[i] & 0xffff;
        v2c[i] = (c == kNoResult || (m.v2c_map[i] >> 16) != 0) ? -1 : c;
      }
      rc = wk_table_upload(dev, m.table, m.ntable / uint64_t(m.col_num), uint32_t(m.col_num), &t);
      if (rc == WK_OK) rc = continue_patterns(dev, m, v2c, &t);
      result = t;
    }
  }
  if (rc && is_device_error(rc)) {
    if (t) (void)wk_table_free(dev, t);
    return rc;
  }
  if (rc) status = rc;

  wk_sparql_msg& a = r.m;
  a.pattern_step = int32_t(m.npatterns);
  if (!status && result) {
    const uint64_t rows = result->rows;
    const uint32_t hcols = result->cols;
    a.gpu_nelems = rows * hcols;
    a.nvars = m.nvars;
    r.v2c.assign(v2c.size(), kNoResult);
    for (size_t i = 0; i < v2c.size(); i++)
      if (v2c[i] >= 0) r.v2c[i] = v2c[i];  // col2ext(col, SID_t)
    if (rows > uint64_t(INT32_MAX)) {
      status = WK_E_OVERFLOW;  // Result::row_num is an int
    } else if (m.blind || rows == 0) {
      // final_process returns at once (sparql.hpp:1420-1421); the table is
      // only fetched when it is not blind
      a.col_num = int32_t(hcols);
      a.row_num = int32_t(rows);
      r.table.clear();
      if (!m.blind && rows) {
        r.table.resize(rows * hcols);
        rc = wk_table_download(dev, result, r.table.data(), r.table.size());
      }
    } else {
      wk_table* fin = nullptr;
      rc = wk_table_final_process(dev, result, v2c.data(), int32_t(v2c.size()), m.required_vars, m.nrequired,
                                  m.distinct, m.order_vars, m.order_desc, m.norders, int64_t(m.offset),
                                  int64_t(m.limit), &fin);
      if (rc == WK_OK) {
        r.table.resize(fin->rows * fin->cols);
        rc = wk_table_download(dev, fin, r.table.data(), r.table.size());
        a.col_num = int32_t(fin->cols);
        a.row_num = int32_t(fin->rows);
        (void)wk_table_free(dev, fin);
      }
    }
    a.gpu_col_num = a.col_num;  // Result::set_col_num under USE_GPU
    if (rc && is_device_error(rc)) {
      if (t) (void)wk_table_free(dev, t);
      return rc;
    }
    if (rc) status = rc;
  }
  if (t) (void)wk_table_free(dev, t);
  if (status) {
    a.col_num = 0;
    a.row_num = 0;
    a.gpu_nelems = 0;
    a.gpu_col_num = 0;
    r.table.clear();
  }
  a.status_code = status;
  // send_reply (gpu_agent.hpp:115-125)
  a.state = WK_SQ_REPLY;
  a.job_type = 0;  // FULL_JOB
  if (m.blind) {   // SPARQLQuery::shrink (query.hpp:775-786)
    r.patterns.clear();
    r.order_vars.clear();
    r.order_desc.clear();
    r.table.clear();
    r.required.clear();
  }
  r.point();
  Out out;
  encode(a, library_version(o), o && o->use_gpu, out);
  dev->agent_reply.swap(out.b);
  *reply_len = dev->agent_reply.size();
  return WK_OK;
}

int wk_agent_reply(wk_dev* dev, uint8_t* dst, uint64_t cap) {
  if (!dev) return WK_E_INVALID_ARG;
  if (dev->agent_reply.empty()) return WK_E_STATE;
  if (!dst || cap < dev->agent_reply.size()) return WK_E_INVALID_ARG;
  std::memcpy(dst, dev->agent_reply.data(), dev->agent_reply.size());
  return WK_OK;
}

}  // extern "C"
