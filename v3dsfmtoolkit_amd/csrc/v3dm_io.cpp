// v3dm_io.cpp -- the host-side C ABI declared in include/v3dm_io.h: SIFT
// feature files, pairwise-match records, the SQLite match database and the
// drivers' bookkeeping.  No GPU call is made here.
//
// Byte layouts (all little-endian, as the reference writes on x86):
//   feature file  "V3DBIN" | uint32 count | count x record
//   record        int32 id | f32 pos[2] | f32 dir[2] | f32 scale | f32 magnitude | f32 desc[128]   (540 B)
//     (V3D/include/Base/v3d_feature.h:41-49 through BinaryOStreamArchive,
//      V3D/include/Base/v3d_serialization.h:335-421, which writes no tags)
//   PairwiseMatch int32 view0, view1 | int32 3, 3, f64 R[9] row-major | int32 3, f64 T[3]
//                 | uint32 n | n x (f32 x, f32 y, int32 view, int32 id) x 2
//     (SfM/reconstruction/include/reconstruction_common.h:65-119 through BlobOArchive;
//      InlineMatrixBase / InlineVectorBase serialise their dimensions first,
//      V3D/include/Math/v3d_linearbase.h:45-52,211-220)
#include <dlfcn.h>

#include <algorithm>
#include <cerrno>
#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>
#include <tuple>
#include <type_traits>
#include <vector>

#include "../../include/v3dm_io.h"

namespace v3dm_host {
extern thread_local std::string g_createError;
}
using v3dm_host::g_createError;

namespace {

constexpr int kDim = 128;
constexpr size_t kRecordBytes = 4 + 6 * 4 + kDim * 4; // 540
constexpr size_t kMagicBytes = 6;

int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
int fail(int code, const char *fmt, ...)
{
   char buf[768];
   va_list ap;
   va_start(ap, fmt);
   vsnprintf(buf, sizeof buf, fmt, ap);
   va_end(ap);
   g_createError = buf;
   return code;
}

int ok()
{
   g_createError.clear();
   return V3DM_OK;
}

// Whole file into memory (feature files are a few MB).
int slurp(const char *path, std::string &out)
{
   if (!path) return fail(V3DM_ERR_ARG, "null path");
   FILE *f = std::fopen(path, "rb");
   if (!f) return fail(V3DM_ERR_ARG, "cannot open %s: %s", path, std::strerror(errno));
   std::fseek(f, 0, SEEK_END);
   const long len = std::ftell(f);
   std::fseek(f, 0, SEEK_SET);
   if (len < 0)
   {
      std::fclose(f);
      return fail(V3DM_ERR_ARG, "cannot size %s", path);
   }
   out.resize((size_t)len);
   const size_t got = len ? std::fread(&out[0], 1, (size_t)len, f) : 0;
   std::fclose(f);
   if (got != (size_t)len) return fail(V3DM_ERR_ARG, "short read of %s", path);
   return V3DM_OK;
}

int magicFormat(const std::string &data, const char *path, int &format)
{
   if (data.size() >= kMagicBytes && std::memcmp(data.data(), "V3DBIN", kMagicBytes) == 0)
      format = V3DM_FEATURES_BINARY;
   else if (data.size() >= kMagicBytes && std::memcmp(data.data(), "V3DTXT", kMagicBytes) == 0)
      format = V3DM_FEATURES_TEXT;
   else
      return fail(V3DM_ERR_ARG, "%s: unknown archive magic tag", path);
   return V3DM_OK;
}

uint32_t getU32(const unsigned char *p) { return (uint32_t)p[0] | (uint32_t)p[1] << 8 | (uint32_t)p[2] << 16 | (uint32_t)p[3] << 24; }
int32_t getI32(const unsigned char *p) { return (int32_t)getU32(p); }
float getF32(const unsigned char *p)
{
   float v;
   std::memcpy(&v, p, 4);
   return v;
}
double getF64(const unsigned char *p)
{
   double v;
   std::memcpy(&v, p, 8);
   return v;
}
void putU32(std::vector<unsigned char> &b, uint32_t v)
{
   for (int i = 0; i < 4; ++i) b.push_back((unsigned char)(v >> (8 * i)));
}
void putI32(std::vector<unsigned char> &b, int32_t v) { putU32(b, (uint32_t)v); }
template <typename T> void putRaw(std::vector<unsigned char> &b, T v)
{
   unsigned char tmp[sizeof(T)];
   std::memcpy(tmp, &v, sizeof(T));
   b.insert(b.end(), tmp, tmp + sizeof(T));
}

// Output of one record, any target may be null.
struct FeatureSink
{
   int32_t *ids;
   float *pos, *dir, *scale, *mag, *desc;
   void put(size_t k, int32_t id, const float v6[6], const float *d) const
   {
      if (ids) ids[k] = id;
      if (pos)
      {
         pos[2 * k] = v6[0];
         pos[2 * k + 1] = v6[1];
      }
      if (dir)
      {
         dir[2 * k] = v6[2];
         dir[2 * k + 1] = v6[3];
      }
      if (scale) scale[k] = v6[4];
      if (mag) mag[k] = v6[5];
      if (desc) std::memcpy(desc + k * kDim, d, kDim * sizeof(float));
   }
};

int binaryCount(const std::string &data, const char *path, uint32_t &count)
{
   if (data.size() < kMagicBytes + 4) return fail(V3DM_ERR_ARG, "%s: truncated before the record count", path);
   count = getU32(reinterpret_cast<const unsigned char *>(data.data()) + kMagicBytes);
   if (count > (uint32_t)INT32_MAX) return fail(V3DM_ERR_ARG, "%s: record count %u too large", path, count);
   return V3DM_OK;
}

int readBinary(const std::string &data, const char *path, int cap, const FeatureSink &sink, int *count)
{
   uint32_t n;
   int rc = binaryCount(data, path, n);
   if (rc) return rc;
   const size_t need = kMagicBytes + 4 + (size_t)n * kRecordBytes;
   if (data.size() < need)
      return fail(V3DM_ERR_ARG, "%s: %u records need %zu bytes, the file has %zu", path, n, need, data.size());
   *count = (int)n;
   if ((int)n > cap) return fail(V3DM_ERR_CAPACITY, "%s holds %u features, the arrays %d", path, n, cap);
   const unsigned char *p = reinterpret_cast<const unsigned char *>(data.data()) + kMagicBytes + 4;
   float v6[6], d[kDim];
   for (uint32_t k = 0; k < n; ++k, p += kRecordBytes)
   {
      const int32_t id = getI32(p);
      for (int i = 0; i < 6; ++i) v6[i] = getF32(p + 4 + 4 * i);
      std::memcpy(d, p + 28, sizeof d);
      sink.put(k, id, v6, d);
   }
   return ok();
}

// Whitespace-separated tokens after the magic, as TextIStreamArchive reads them.
struct Tokens
{
   const char *p, *end;
   const char *path;
   bool next(std::string &tok)
   {
      while (p < end && (*p == ' ' || *p == '\n' || *p == '\t' || *p == '\r' || *p == '\v' || *p == '\f')) ++p;
      if (p >= end) return false;
      const char *s = p;
      while (p < end && !(*p == ' ' || *p == '\n' || *p == '\t' || *p == '\r' || *p == '\v' || *p == '\f')) ++p;
      tok.assign(s, p);
      return true;
   }
   int expect(const char *want, const char *what)
   {
      std::string t;
      if (!next(t)) return fail(V3DM_ERR_ARG, "%s: text archive ends before %s", path, what);
      if (t != want) return fail(V3DM_ERR_ARG, "%s: tag mismatch <%s> instead of <%s>", path, t.c_str(), want);
      return V3DM_OK;
   }
   int integer(long long lo, long long hi, long long &v, const char *what)
   {
      std::string t;
      if (!next(t)) return fail(V3DM_ERR_ARG, "%s: text archive ends before %s", path, what);
      char *e = nullptr;
      errno = 0;
      v = std::strtoll(t.c_str(), &e, 10);
      if (errno || *e || e == t.c_str() || v < lo || v > hi)
         return fail(V3DM_ERR_ARG, "%s: bad %s <%s>", path, what, t.c_str());
      return V3DM_OK;
   }
   int real(float &v, const char *what)
   {
      std::string t;
      if (!next(t)) return fail(V3DM_ERR_ARG, "%s: text archive ends before %s", path, what);
      char *e = nullptr;
      v = std::strtof(t.c_str(), &e);
      if (*e || e == t.c_str()) return fail(V3DM_ERR_ARG, "%s: bad %s <%s>", path, what, t.c_str());
      return V3DM_OK;
   }
};

int textCount(const std::string &data, const char *path, Tokens &tk, uint32_t &n)
{
   tk = Tokens{data.data() + kMagicBytes, data.data() + data.size(), path};
   long long v;
   int rc = tk.integer(0, INT32_MAX, v, "record count");
   if (rc) return rc;
   n = (uint32_t)v;
   return V3DM_OK;
}

int readText(const std::string &data, const char *path, int cap, const FeatureSink &sink, int *count)
{
   Tokens tk;
   uint32_t n;
   int rc = textCount(data, path, tk, n);
   if (rc) return rc;
   *count = (int)n;
   if ((int)n > cap) return fail(V3DM_ERR_CAPACITY, "%s holds %u features, the arrays %d", path, n, cap);
   if ((rc = tk.expect("{", "the vector scope"))) return rc;
   float v6[6], d[kDim];
   for (uint32_t k = 0; k < n; ++k)
   {
      if ((rc = tk.expect("SIFT_Feature", "a feature tag")) || (rc = tk.expect("{", "a feature scope"))) return rc;
      long long id;
      if ((rc = tk.integer(INT32_MIN, INT32_MAX, id, "feature id"))) return rc;
      for (int i = 0; i < 6; ++i)
         if ((rc = tk.real(v6[i], "feature value"))) return rc;
      for (int i = 0; i < kDim; ++i)
         if ((rc = tk.real(d[i], "descriptor value"))) return rc;
      if ((rc = tk.expect("}", "the end of a feature"))) return rc;
      sink.put(k, (int32_t)id, v6, d);
   }
   if ((rc = tk.expect("}", "the end of the vector"))) return rc;
   return ok();
}

// ---- SQLite, loaded at run time (the image has the library, not its header) ----

struct Sqlite
{
   void *h = nullptr;
   int (*open)(const char *, void **) = nullptr;
   int (*close)(void *) = nullptr;
   int (*exec)(void *, const char *, int (*)(void *, int, char **, char **), void *, char **) = nullptr;
   int (*prepare)(void *, const char *, int, void **, const char **) = nullptr;
   int (*bindInt)(void *, int, int) = nullptr;
   int (*bindBlob)(void *, int, const void *, int, void (*)(void *)) = nullptr;
   int (*step)(void *) = nullptr;
   int (*reset)(void *) = nullptr;
   int (*finalize)(void *) = nullptr;
   const char *(*errmsg)(void *) = nullptr;
   void (*freeMem)(void *) = nullptr;
};

constexpr int kSqliteOk = 0;
constexpr int kSqliteDone = 101;

void loadSqlite(Sqlite &s)
{
   void *h = dlopen("libsqlite3.so.0", RTLD_NOW | RTLD_LOCAL);
   if (!h) return;
   bool all = true;
   auto sym = [&](auto &fn, const char *name) {
      fn = reinterpret_cast<std::remove_reference_t<decltype(fn)>>(dlsym(h, name));
      all = all && fn;
   };
   sym(s.open, "sqlite3_open");
   sym(s.close, "sqlite3_close");
   sym(s.exec, "sqlite3_exec");
   sym(s.prepare, "sqlite3_prepare_v2");
   sym(s.bindInt, "sqlite3_bind_int");
   sym(s.bindBlob, "sqlite3_bind_blob");
   sym(s.step, "sqlite3_step");
   sym(s.reset, "sqlite3_reset");
   sym(s.finalize, "sqlite3_finalize");
   sym(s.errmsg, "sqlite3_errmsg");
   sym(s.freeMem, "sqlite3_free");
   if (!all)
   {
      dlclose(h);
      return;
   }
   s.h = h;
}

// The library, loaded once per process (thread-safe); null if unavailable.
const Sqlite *sqlite()
{
   static Sqlite s;
   static std::once_flag once;
   std::call_once(once, [] { loadSqlite(s); });
   return s.h ? &s : nullptr;
}

void encodeMatch(std::vector<unsigned char> &b, int view0, int view1, const double *R, const double *T, int n,
                 const v3dm_correspondence *corrs)
{
   b.clear();
   b.reserve(v3dm_pairwise_match_size(n));
   putI32(b, view0);
   putI32(b, view1);
   putI32(b, 3);
   putI32(b, 3);
   for (int i = 0; i < 9; ++i) putRaw(b, R[i]);
   putI32(b, 3);
   for (int i = 0; i < 3; ++i) putRaw(b, T[i]);
   putU32(b, (uint32_t)n);
   for (int k = 0; k < n; ++k)
      for (const v3dm_point_measurement *m : {&corrs[k].left, &corrs[k].right})
      {
         putRaw(b, m->x);
         putRaw(b, m->y);
         putI32(b, m->view);
         putI32(b, m->id);
      }
}

} // namespace

static_assert(sizeof(v3dm_point_measurement) == 16, "PointMeasurement is 16 bytes");
static_assert(sizeof(v3dm_correspondence) == 32, "PointCorrespondence is 32 bytes");

struct v3dm_match_db
{
   void *db = nullptr;
   void *putData = nullptr, *putList = nullptr;
   std::vector<unsigned char> blob;
};

extern "C" {

int v3dm_features_info(const char *path, int *count, int *format)
{
   if (!count || !format) return fail(V3DM_ERR_ARG, "null output");
   std::string data;
   int rc = slurp(path, data);
   if (rc) return rc;
   if ((rc = magicFormat(data, path, *format))) return rc;
   uint32_t n;
   if (*format == V3DM_FEATURES_BINARY)
      rc = binaryCount(data, path, n);
   else
   {
      Tokens tk;
      rc = textCount(data, path, tk, n);
   }
   if (rc) return rc;
   *count = (int)n;
   return ok();
}

int v3dm_features_read(const char *path, int cap, int32_t *ids, float *positions, float *directions, float *scales,
                       float *magnitudes, float *descriptors, int *count)
{
   if (!count || cap < 0) return fail(V3DM_ERR_ARG, "null count or negative capacity");
   *count = 0;
   std::string data;
   int rc = slurp(path, data);
   if (rc) return rc;
   int format;
   if ((rc = magicFormat(data, path, format))) return rc;
   const FeatureSink sink{ids, positions, directions, scales, magnitudes, descriptors};
   return format == V3DM_FEATURES_BINARY ? readBinary(data, path, cap, sink, count)
                                         : readText(data, path, cap, sink, count);
}

int v3dm_features_write(const char *path, int n, const int32_t *ids, const float *positions, const float *directions,
                        const float *scales, const float *magnitudes, const float *descriptors, int format)
{
   if (!path || n < 0 || (n > 0 && !descriptors)) return fail(V3DM_ERR_ARG, "null path/descriptors or negative n");
   if (format != V3DM_FEATURES_BINARY && format != V3DM_FEATURES_TEXT)
      return fail(V3DM_ERR_ARG, "unknown feature-file format %d", format);
   auto value = [&](size_t k, int i) -> float {
      switch (i)
      {
      case 0:
      case 1: return positions ? positions[2 * k + i] : 0.0f;
      case 2:
      case 3: return directions ? directions[2 * k + i - 2] : 0.0f;
      case 4: return scales ? scales[k] : 0.0f;
      default: return magnitudes ? magnitudes[k] : 0.0f;
      }
   };
   std::string out;
   if (format == V3DM_FEATURES_BINARY)
   {
      std::vector<unsigned char> b;
      b.reserve(kMagicBytes + 4 + (size_t)n * kRecordBytes);
      b.insert(b.end(), "V3DBIN", "V3DBIN" + kMagicBytes);
      putU32(b, (uint32_t)n);
      for (size_t k = 0; k < (size_t)n; ++k)
      {
         putI32(b, ids ? ids[k] : (int32_t)k);
         for (int i = 0; i < 6; ++i) putRaw(b, value(k, i));
         const unsigned char *d = reinterpret_cast<const unsigned char *>(descriptors + k * kDim);
         b.insert(b.end(), d, d + kDim * sizeof(float));
      }
      out.assign(b.begin(), b.end());
   }
   else
   {
      // TextOStreamArchive (v3d_serialization.h:173-256): every value followed by
      // one blank; a scope opens with "{ " and closes with "} ", a newline and
      // the indentation of the enclosing scope (two blanks per level).
      std::ostringstream os;
      os << "V3DTXT" << '\n';
      os << (unsigned)n << " { ";
      for (size_t k = 0; k < (size_t)n; ++k)
      {
         os << "SIFT_Feature { " << (ids ? ids[k] : (int32_t)k) << ' ';
         for (int i = 0; i < 6; ++i) os << value(k, i) << ' ';
         for (int i = 0; i < kDim; ++i) os << descriptors[k * kDim + i] << ' ';
         os << "} \n  ";
      }
      os << "} \n";
      out = os.str();
   }
   FILE *f = std::fopen(path, "wb");
   if (!f) return fail(V3DM_ERR_ARG, "cannot create %s: %s", path, std::strerror(errno));
   const size_t put = out.empty() ? 0 : std::fwrite(out.data(), 1, out.size(), f);
   const bool closed = std::fclose(f) == 0;
   if (put != out.size() || !closed) return fail(V3DM_ERR_ARG, "short write to %s", path);
   return ok();
}

int v3dm_truncated_feature_count(int64_t size, int nMax)
{
   const int t = size <= 0 ? 0 : (int)((uint64_t)size & 0xffe0u);
   return std::min(nMax, t);
}

int v3dm_sequence_pairs(int nViews, int radius, int32_t *pairs, int cap, int *count)
{
   if (!count || nViews < 0 || cap < 0 || (cap > 0 && !pairs)) return fail(V3DM_ERR_ARG, "bad arguments");
   int64_t n = 0;
   for (int i = 0; i < nViews; ++i)
      for (int64_t j = (int64_t)i + 1; j < (int64_t)i + 2 * (int64_t)radius; ++j)
      {
         if (j >= nViews) continue;
         if (n < cap)
         {
            pairs[2 * n] = i;
            pairs[2 * n + 1] = (int32_t)j;
         }
         ++n;
      }
   if (n > INT32_MAX) return fail(V3DM_ERR_ARG, "too many pairs");
   *count = (int)n;
   if (n > cap) return fail(V3DM_ERR_CAPACITY, "%lld pairs, capacity %d", (long long)n, cap);
   return ok();
}

int v3dm_descriptors_to_u8(int64_t n, const float *descriptors, uint8_t *out)
{
   if (n < 0 || (n > 0 && (!descriptors || !out))) return fail(V3DM_ERR_ARG, "bad arguments");
   const int64_t total = n * kDim;
   for (int64_t i = 0; i < total; ++i)
   {
      const float s = descriptors[i] * 512.0f; // exact: a power-of-two scale
      if (!(s >= 0.0f && s <= 255.0f) || s != std::floor(s))
         return fail(V3DM_ERR_UNSUPPORTED, "descriptor value %d of row %lld is not u/512", (int)(i % kDim),
                     (long long)(i / kDim));
      out[i] = (uint8_t)s;
   }
   return ok();
}

size_t v3dm_pairwise_match_size(int nCorrs) { return 8 + 80 + 28 + 4 + 32 * (size_t)(nCorrs < 0 ? 0 : nCorrs); }

int v3dm_encode_pairwise_match(int view0, int view1, const double rotation[9], const double translation[3], int nCorrs,
                               const v3dm_correspondence *corrs, uint8_t *blob, size_t cap, size_t *size)
{
   if (!rotation || !translation || !size || nCorrs < 0 || (nCorrs > 0 && !corrs))
      return fail(V3DM_ERR_ARG, "bad arguments");
   *size = v3dm_pairwise_match_size(nCorrs);
   if (cap < *size || !blob) return fail(V3DM_ERR_CAPACITY, "blob needs %zu bytes, capacity %zu", *size, cap);
   std::vector<unsigned char> b;
   encodeMatch(b, view0, view1, rotation, translation, nCorrs, corrs);
   std::memcpy(blob, b.data(), b.size());
   return ok();
}

int v3dm_decode_pairwise_match(const uint8_t *blob, size_t size, int *view0, int *view1, double rotation[9],
                               double translation[3], v3dm_correspondence *corrs, int cap, int *nCorrs)
{
   if (!blob || !view0 || !view1 || !rotation || !translation || !nCorrs || cap < 0)
      return fail(V3DM_ERR_ARG, "bad arguments");
   if (size < v3dm_pairwise_match_size(0)) return fail(V3DM_ERR_ARG, "blob of %zu bytes is too short", size);
   const unsigned char *p = blob;
   *view0 = getI32(p);
   *view1 = getI32(p + 4);
   if (getI32(p + 8) != 3 || getI32(p + 12) != 3) return fail(V3DM_ERR_ARG, "rotation is not 3 x 3");
   for (int i = 0; i < 9; ++i) rotation[i] = getF64(p + 16 + 8 * i);
   if (getI32(p + 88) != 3) return fail(V3DM_ERR_ARG, "translation is not 3-D");
   for (int i = 0; i < 3; ++i) translation[i] = getF64(p + 92 + 8 * i);
   const uint32_t n = getU32(p + 116);
   if (n > (uint32_t)INT32_MAX || size != v3dm_pairwise_match_size((int)n))
      return fail(V3DM_ERR_ARG, "blob of %zu bytes does not hold %u correspondences", size, n);
   *nCorrs = (int)n;
   if ((int)n > cap || (n > 0 && !corrs)) return fail(V3DM_ERR_CAPACITY, "%u correspondences, capacity %d", n, cap);
   const unsigned char *q = p + 120;
   for (uint32_t k = 0; k < n; ++k)
      for (v3dm_point_measurement *m : {&corrs[k].left, &corrs[k].right})
      {
         m->x = getF32(q);
         m->y = getF32(q + 4);
         m->view = getI32(q + 8);
         m->id = getI32(q + 12);
         q += 16;
      }
   return ok();
}

int v3dm_match_db_open(const char *path, int truncate, v3dm_match_db **out)
{
   if (!path || !out) return fail(V3DM_ERR_ARG, "null path or output");
   *out = nullptr;
   const Sqlite *s = sqlite();
   if (!s) return fail(V3DM_ERR_UNSUPPORTED, "libsqlite3.so.0 could not be loaded");
   auto db = std::make_unique<v3dm_match_db>();
   if (s->open(path, &db->db) != kSqliteOk)
   {
      std::string m = db->db ? s->errmsg(db->db) : "sqlite3_open failed";
      if (db->db) s->close(db->db);
      return fail(V3DM_ERR_ARG, "cannot open %s: %s", path, m.c_str());
   }
   auto run = [&](const char *sql) -> int {
      char *msg = nullptr;
      if (s->exec(db->db, sql, nullptr, nullptr, &msg) != kSqliteOk)
      {
         std::string m = msg ? msg : "?";
         if (msg) s->freeMem(msg);
         return fail(V3DM_ERR_ARG, "SQL error (%s): %s", sql, m.c_str());
      }
      return V3DM_OK;
   };
   int rc = run("pragma synchronous = off");
   for (const char *t : {"matches_data", "matches_list"})
   {
      char sql[128];
      if (!rc && truncate)
      {
         std::snprintf(sql, sizeof sql, "drop table if exists %s", t);
         rc = run(sql);
      }
      if (!rc)
      {
         std::snprintf(sql, sizeof sql, "create table if not exists %s (DATA BLOB)", t);
         rc = run(sql);
      }
   }
   if (!rc) rc = run("begin");
   if (!rc)
      for (auto [stmt, t] : {std::pair<void **, const char *>{&db->putData, "matches_data"},
                             std::pair<void **, const char *>{&db->putList, "matches_list"}})
      {
         char sql[128];
         std::snprintf(sql, sizeof sql, "insert or replace into %s (OID, DATA) values (?, ?)", t);
         if (s->prepare(db->db, sql, -1, stmt, nullptr) != kSqliteOk)
         {
            rc = fail(V3DM_ERR_ARG, "SQL error (%s): %s", sql, s->errmsg(db->db));
            break;
         }
      }
   if (rc)
   {
      if (db->putData) s->finalize(db->putData);
      if (db->putList) s->finalize(db->putList);
      s->close(db->db);
      return rc;
   }
   *out = db.release();
   return ok();
}

int v3dm_match_db_put(v3dm_match_db *db, int oid, int view0, int view1, const double rotation[9],
                      const double translation[3], int nCorrs, const v3dm_correspondence *corrs)
{
   if (!db || !rotation || !translation || nCorrs < 0 || (nCorrs > 0 && !corrs))
      return fail(V3DM_ERR_ARG, "bad arguments");
   const Sqlite *s = sqlite();
   encodeMatch(db->blob, view0, view1, rotation, translation, nCorrs, corrs);
   if (db->blob.size() > (size_t)INT32_MAX) return fail(V3DM_ERR_ARG, "record too large");
   unsigned char vp[8];
   for (int i = 0; i < 4; ++i)
   {
      vp[i] = (unsigned char)((uint32_t)view0 >> (8 * i));
      vp[4 + i] = (unsigned char)((uint32_t)view1 >> (8 * i));
   }
   void (*transient)(void *) = reinterpret_cast<void (*)(void *)>(intptr_t(-1)); // SQLITE_TRANSIENT
   for (auto [stmt, data, len] : {std::tuple<void *, const void *, int>{db->putData, db->blob.data(), (int)db->blob.size()},
                                  std::tuple<void *, const void *, int>{db->putList, vp, 8}})
   {
      s->bindInt(stmt, 1, oid);
      s->bindBlob(stmt, 2, data, len, transient);
      const int res = s->step(stmt);
      s->reset(stmt);
      if (res != kSqliteDone) return fail(V3DM_ERR_ARG, "SQL insert failed: %s", s->errmsg(db->db));
   }
   return ok();
}

int v3dm_match_db_close(v3dm_match_db *db)
{
   if (!db) return ok();
   const Sqlite *s = sqlite();
   s->finalize(db->putData);
   s->finalize(db->putList);
   char *msg = nullptr;
   int rc = V3DM_OK;
   if (s->exec(db->db, "commit", nullptr, nullptr, &msg) != kSqliteOk)
   {
      rc = fail(V3DM_ERR_ARG, "SQL commit failed: %s", msg ? msg : "?");
      if (msg) s->freeMem(msg);
   }
   s->close(db->db);
   delete db;
   return rc ? rc : ok();
}

} // extern "C"
