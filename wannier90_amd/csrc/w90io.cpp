// w90io.cpp — readers of the .mmn and .amn overlap files on the host, in
// parallel over the file.
//
// The records are those of overlap_read (src/overlap.F90:234-368): after a
// comment line and a line of sizes, the .mmn holds num_kpts*nntot blocks,
// each a header "k k' G1 G2 G3" followed by num_bands^2 records "Re Im"
// (row index fastest); the .amn holds num_bands*num_proj*num_kpts records
// "m n k Re Im".  Every record is one non-blank line; blank lines are
// skipped, values are separated by blanks, tabs or commas, and what follows
// the values a record needs is ignored, as a list-directed read does.  (A
// list-directed read would also continue a record whose line holds too few
// values on the next line; here such a line is a read error.)
//
// The file is mapped and cut into one byte range per thread at line ends.
// A first pass counts the records of each range, so that the record number
// of every line is known; the second parses each range in place: a .mmn
// block goes to the b with nnlist(k,b) = k' and nncell(:,k,b) = G, found
// from its header.  The elements at the start of a range whose block header
// lies in the range before are parsed after all ranges have read their
// headers.  When two blocks (or two .amn records) name the same
// destination the later one wins, as in the sequential read: those
// destinations are rewritten from the later record at the end.
//
// Decimal-to-binary conversion is correctly rounded: a number with at most
// 19 significant digits whose integer mantissa w <= 2^53 and decimal
// exponent |e| <= 22 is w * 10^e or w / 10^-e, one IEEE operation on exact
// operands (Clinger's fast path, which covers the f18.12 records that
// pw2wannier90 writes); any other number goes to strtod_l in the C locale.
// Fortran exponent letters d/D and exponents written without a letter
// ("1.5-003") are accepted.

#include "w90gpu.h"

#include <fcntl.h>
#include <locale.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

namespace {

void set_msg(char* errmsg, int errlen, const std::string& s) {
  if (!errmsg || errlen <= 0) return;
  const size_t n = std::min(s.size(), (size_t)errlen - 1);
  std::memcpy(errmsg, s.data(), n);
  errmsg[n] = '\0';
}

// Read-only mapping of a whole file.
struct Mapped {
  const char* p = nullptr;
  size_t n = 0;
  int fd = -1;
  bool open(const char* path) {
    fd = ::open(path, O_RDONLY);
    if (fd < 0) return false;
    struct stat sb;
    if (fstat(fd, &sb) != 0) return false;
    n = (size_t)sb.st_size;
    if (n == 0) return true;
    void* q = mmap(nullptr, n, PROT_READ, MAP_PRIVATE, fd, 0);
    if (q == MAP_FAILED) {
      n = 0;
      return false;
    }
    p = static_cast<const char*>(q);
    return true;
  }
  ~Mapped() {
    if (p) munmap(const_cast<char*>(p), n);
    if (fd >= 0) ::close(fd);
  }
};

inline bool is_blank(char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\f' || c == '\v'; }
inline bool is_sep(char c) { return is_blank(c) || c == ','; }
inline bool is_digit(char c) { return c >= '0' && c <= '9'; }

inline const char* line_end(const char* s, const char* e) {
  const void* q = std::memchr(s, '\n', (size_t)(e - s));
  return q ? static_cast<const char*>(q) : e;
}

inline bool blank_line(const char* s, const char* le) {
  while (s < le && is_blank(*s)) ++s;
  return s == le;
}

// Next value of a record: separators are skipped, [*b, *t) is the token.
inline bool next_token(const char*& s, const char* le, const char** b, const char** t) {
  while (s < le && is_sep(*s)) ++s;
  if (s >= le) return false;
  *b = s;
  while (s < le && !is_sep(*s)) ++s;
  *t = s;
  return true;
}

bool parse_int(const char* b, const char* t, long* v) {
  bool neg = false;
  if (b < t && (*b == '+' || *b == '-')) {
    neg = *b == '-';
    ++b;
  }
  if (b == t) return false;
  long x = 0;
  for (; b < t; ++b) {
    if (!is_digit(*b) || x > 100000000000L) return false;
    x = x * 10 + (*b - '0');
  }
  *v = neg ? -x : x;
  return true;
}

const double kPow10[23] = {1e0,  1e1,  1e2,  1e3,  1e4,  1e5,  1e6,  1e7,  1e8,  1e9,  1e10, 1e11,
                           1e12, 1e13, 1e14, 1e15, 1e16, 1e17, 1e18, 1e19, 1e20, 1e21, 1e22};

locale_t c_locale() {
  static locale_t loc = newlocale(LC_ALL_MASK, "C", (locale_t)0);
  return loc;
}

// Any number the fast path leaves: Fortran forms rewritten for strtod_l.
bool parse_real_slow(const char* b, const char* t, double* v) {
  char buf[128];
  size_t n = 0;
  const size_t len = (size_t)(t - b);
  if (len + 2 > sizeof(buf)) return false;
  bool seen_exp = false;
  for (const char* s = b; s < t; ++s) {
    char c = *s;
    if (c == 'd' || c == 'D' || c == 'e' || c == 'E') {
      c = 'e';
      seen_exp = true;
    } else if ((c == '+' || c == '-') && s > b && !seen_exp && (is_digit(s[-1]) || s[-1] == '.')) {
      buf[n++] = 'e';  // exponent without a letter
      seen_exp = true;
    }
    buf[n++] = c;
  }
  buf[n] = '\0';
  char* end = nullptr;
  const double x = strtod_l(buf, &end, c_locale());
  if (end != buf + n) return false;
  *v = x;
  return true;
}

inline bool parse_real(const char* b, const char* t, double* v) {
  const char* s = b;
  bool neg = false;
  if (s < t && (*s == '+' || *s == '-')) {
    neg = *s == '-';
    ++s;
  }
  uint64_t w = 0;
  int nd = 0, dexp = 0;
  bool any = false, exact = true;
  for (; s < t && is_digit(*s); ++s) {
    any = true;
    const int d = *s - '0';
    if (w == 0 && d == 0) continue;
    if (nd < 19) {
      w = w * 10 + (uint64_t)d;
      ++nd;
    } else {
      exact = false;
    }
  }
  if (s < t && *s == '.') {
    for (++s; s < t && is_digit(*s); ++s) {
      any = true;
      const int d = *s - '0';
      if (w == 0 && d == 0) {
        --dexp;
        continue;
      }
      if (nd < 19) {
        w = w * 10 + (uint64_t)d;
        ++nd;
        --dexp;
      } else {
        exact = false;
      }
    }
  }
  if (!any) return parse_real_slow(b, t, v);
  if (s < t) {
    const char c = *s;
    if (c == 'e' || c == 'E' || c == 'd' || c == 'D') ++s;
    else if (c != '+' && c != '-') return parse_real_slow(b, t, v);
    bool eneg = false;
    if (s < t && (*s == '+' || *s == '-')) {
      eneg = *s == '-';
      ++s;
    }
    if (s == t) return parse_real_slow(b, t, v);
    int ex = 0;
    for (; s < t; ++s) {
      if (!is_digit(*s)) return parse_real_slow(b, t, v);
      if (ex < 100000) ex = ex * 10 + (*s - '0');
    }
    dexp += eneg ? -ex : ex;
  }
  if (w == 0) {
    *v = neg ? -0.0 : 0.0;
    return true;
  }
  if (exact && w <= (uint64_t(1) << 53) && dexp >= -22 && dexp <= 22) {
    double x = (double)w;
    x = dexp >= 0 ? x * kPow10[dexp] : x / kPow10[-dexp];
    *v = neg ? -x : x;
    return true;
  }
  return parse_real_slow(b, t, v);
}

// nthreads > 0 is taken as given (at most 64); otherwise OMP_NUM_THREADS or
// the hardware threads, with at least 1 MiB of text per thread.
int pick_threads(int nthreads, size_t bytes) {
  if (nthreads > 0) return std::min(nthreads, 64);
  const char* omp = std::getenv("OMP_NUM_THREADS");
  int t = omp ? std::atoi(omp) : 0;
  if (t <= 0) t = (int)std::thread::hardware_concurrency();
  t = std::max(1, std::min(t, 64));
  const size_t per = 1 << 20;
  return (int)std::min<size_t>((size_t)t, bytes / per + 1);
}

// Byte ranges [cut[i], cut[i+1]) of the body, each starting at a line.
std::vector<const char*> cut_ranges(const char* b, const char* e, int nt) {
  std::vector<const char*> cut(nt + 1);
  cut[0] = b;
  cut[nt] = e;
  const size_t n = (size_t)(e - b);
  for (int i = 1; i < nt; ++i) {
    const char* s = b + n * (size_t)i / (size_t)nt;
    if (s < cut[i - 1]) s = cut[i - 1];
    const char* le = line_end(s, e);
    cut[i] = le < e ? le + 1 : e;
  }
  return cut;
}

// f(i) for every range i, one thread per range; ranges whose thread cannot
// be started run on the calling thread (no exception leaves the C ABI).
template <class F>
void run_threads(int nt, F f) {
  std::vector<std::thread> th;
  int started = 0;
  if (nt > 1) {
    try {
      th.reserve(nt - 1);
      for (int i = 1; i < nt; ++i) {
        th.emplace_back(f, i);
        ++started;
      }
    } catch (...) {
    }
  }
  f(0);
  for (int i = 1 + started; i < nt; ++i) f(i);
  for (auto& x : th) x.join();
}

// Records (non-blank lines) of each range.
std::vector<long> count_records(const std::vector<const char*>& cut) {
  const int nt = (int)cut.size() - 1;
  std::vector<long> cnt(nt, 0);
  run_threads(nt, [&](int i) {
    long c = 0;
    for (const char* s = cut[i]; s < cut[i + 1];) {
      const char* le = line_end(s, cut[i + 1]);
      if (!blank_line(s, le)) ++c;
      s = le + 1;
    }
    cnt[i] = c;
  });
  return cnt;
}

// The first error in record order wins, as in a sequential read.
struct Err {
  long rec = -1;
  std::string msg;
  void set(long r, const std::string& m) {
    if (rec < 0 || r < rec) {
      rec = r;
      msg = m;
    }
  }
};

// First two lines: comment, then three sizes.  Returns the body start.
const char* read_sizes(const Mapped& f, long sz[3], bool* ok) {
  *ok = false;
  const char* e = f.p + f.n;
  const char* s = f.p;
  const char* le = line_end(s, e);
  if (le >= e) return e;
  s = le + 1;
  while (s < e) {  // blank records before the sizes are skipped
    le = line_end(s, e);
    if (!blank_line(s, le)) break;
    s = le + 1;
  }
  if (s >= e) return e;
  const char* q = s;
  for (int i = 0; i < 3; ++i) {
    const char *b, *t;
    if (!next_token(q, le, &b, &t) || !parse_int(b, t, &sz[i])) return e;
  }
  *ok = true;
  return le < e ? le + 1 : e;
}

int read_mmn(const char* path, int num_bands, int num_kpts, int nntot, const int* nnlist, const int* nncell,
             double* m_matrix, int nthreads, char* errmsg, int errlen) {
  set_msg(errmsg, errlen, "");
  if (!path || !nnlist || !nncell || !m_matrix || num_bands <= 0 || num_kpts <= 0 || nntot <= 0) {
    set_msg(errmsg, errlen, "w90gpu_read_mmn: invalid arguments");
    return W90GPU_ERR_INPUT;
  }
  const std::string name(path);
  Mapped f;
  if (!f.open(path)) {
    set_msg(errmsg, errlen, "Error: Problem opening input file " + name);
    return W90GPU_ERR_FILE;
  }
  const std::string read_err = "Error: Problem reading input file " + name;
  long sz[3];
  bool ok = false;
  const char* body = f.n ? read_sizes(f, sz, &ok) : nullptr;
  if (!ok) {
    set_msg(errmsg, errlen, read_err);
    return W90GPU_ERR_FILE;
  }
  if (sz[0] != num_bands) {
    set_msg(errmsg, errlen, name + " has not the right number of bands");
    return W90GPU_ERR_FILE;
  }
  if (sz[1] != num_kpts) {
    set_msg(errmsg, errlen, name + " has not the right number of k-points");
    return W90GPU_ERR_FILE;
  }
  if (sz[2] != nntot) {
    set_msg(errmsg, errlen, name + " has not the right number of nearest neighbours");
    return W90GPU_ERR_FILE;
  }
  const long nb = num_bands, nk = num_kpts, nn = nntot;
  const long nb2 = nb * nb, per_blk = nb2 + 1, nblk = nk * nn;
  const long nrec = nblk * per_blk;
  const char* end = f.p + f.n;
  const int nt = pick_threads(nthreads, (size_t)(end - body));
  const std::vector<const char*> cut = cut_ranges(body, end, nt);
  const std::vector<long> cnt = count_records(cut);
  std::vector<long> r0(nt + 1, 0);
  for (int i = 0; i < nt; ++i) r0[i + 1] = r0[i] + cnt[i];
  if (r0[nt] < nrec) {
    set_msg(errmsg, errlen, read_err);
    return W90GPU_ERR_FILE;
  }

  // destination (k, b) of each block: -1 unknown, -2 error
  std::vector<long> dest(nblk, -1);
  std::vector<const char*> hdr_pos(nblk, nullptr);
  std::vector<Err> errs(nt);
  std::vector<const char*> pend_b(nt, nullptr), pend_e(nt, nullptr);
  auto find_b = [&](long k, long k2, long g0, long g1, long g2, long rec, Err& er) -> long {
    if (k < 1 || k > nk) {
      er.set(rec, "Neighbour not found");
      return -2;
    }
    long b = -1;
    for (long inn = 0; inn < nn; ++inn) {
      const long q = (k - 1) + nk * inn;
      if (nnlist[q] == k2 && nncell[3 * q] == g0 && nncell[3 * q + 1] == g1 && nncell[3 * q + 2] == g2) {
        if (b >= 0) {
          er.set(rec, "Error reading " + name + ". More than one matching nearest neighbour found");
          return -2;
        }
        b = inn;
      }
    }
    if (b < 0) {
      er.set(rec, "Neighbour not found");
      return -2;
    }
    return b + nn * (k - 1);
  };
  // one element record [s, le) with record number r
  auto parse_elem = [&](const char* s, const char* le, long r, Err& er) {
    const long blk = r / per_blk, q = r % per_blk;
    if (q == 0 || dest[blk] < 0) return;
    const char *b1, *t1, *b2, *t2;
    double re, im;
    if (!next_token(s, le, &b1, &t1) || !next_token(s, le, &b2, &t2) || !parse_real(b1, t1, &re) ||
        !parse_real(b2, t2, &im)) {
      er.set(r, read_err);
      return;
    }
    double* out = m_matrix + 2 * (dest[blk] * nb2 + (q - 1));
    out[0] = re;
    out[1] = im;
  };
  // the element records in [s, e), the first with record number r, up to
  // record rend
  auto parse_elems = [&](const char* s, const char* e, long r, long rend, Err& er) {
    while (s < e && r < rend) {
      const char* le = line_end(s, e);
      if (!blank_line(s, le)) parse_elem(s, le, r++, er);
      s = le + 1;
    }
  };
  // headers, and the elements of blocks whose header is in this range
  run_threads(nt, [&](int i) {
    Err& er = errs[i];
    const char* s = cut[i];
    const char* e = cut[i + 1];
    long r = r0[i];
    bool pending = true;  // elements before the first header of the range
    pend_b[i] = s;
    pend_e[i] = s;
    while (s < e && r < nrec) {
      const char* le = line_end(s, e);
      if (blank_line(s, le)) {
        s = le + 1;
        continue;
      }
      const long blk = r / per_blk, q = r % per_blk;
      if (q == 0) {
        if (pending) {
          pend_e[i] = s;
          pending = false;
        }
        long v[5];
        const char* c = s;
        bool good = true;
        for (int j = 0; j < 5 && good; ++j) {
          const char *b, *t;
          good = next_token(c, le, &b, &t) && parse_int(b, t, &v[j]);
        }
        hdr_pos[blk] = s;
        if (!good) {
          er.set(r, read_err);
          dest[blk] = -2;
        } else {
          // the reference checks the neighbour after reading the block
          dest[blk] = find_b(v[0], v[1], v[2], v[3], v[4], r + nb2, er);
        }
      } else if (!pending) {
        parse_elem(s, le, r, er);
      }
      ++r;
      s = le + 1;
    }
    if (pending) pend_e[i] = s;
  });
  // elements whose header was read by the range before
  run_threads(nt, [&](int i) { parse_elems(pend_b[i], pend_e[i], r0[i], nrec, errs[i]); });
  Err first;
  for (const Err& er : errs)
    if (er.rec >= 0) first.set(er.rec, er.msg);
  if (first.rec >= 0) {
    set_msg(errmsg, errlen, first.msg);
    return W90GPU_ERR_FILE;
  }
  // the later of two blocks with one destination wins; destinations no
  // block names are zero
  std::vector<long> last(nblk, -1);
  bool dup = false;
  for (long blk = 0; blk < nblk; ++blk) {
    if (last[dest[blk]] >= 0) dup = true;
    last[dest[blk]] = blk;
  }
  for (long d = 0; d < nblk; ++d)
    if (last[d] < 0) std::memset(m_matrix + 2 * d * nb2, 0, sizeof(double) * 2 * (size_t)nb2);
  if (dup) {
    Err er;
    for (long d = 0; d < nblk; ++d) {
      const long blk = last[d];
      if (blk < 0) continue;
      const char* s = line_end(hdr_pos[blk], end);
      s = s < end ? s + 1 : end;
      parse_elems(s, end, blk * per_blk + 1, (blk + 1) * per_blk, er);
    }
  }
  return W90GPU_OK;
}

int read_amn(const char* path, int num_bands, int num_kpts, int num_proj, double* a_matrix, int nthreads,
             char* errmsg, int errlen) {
  set_msg(errmsg, errlen, "");
  if (!path || !a_matrix || num_bands <= 0 || num_kpts <= 0 || num_proj <= 0) {
    set_msg(errmsg, errlen, "w90gpu_read_amn: invalid arguments");
    return W90GPU_ERR_INPUT;
  }
  const std::string name(path);
  Mapped f;
  if (!f.open(path)) {
    set_msg(errmsg, errlen, "Error: Problem opening input file " + name);
    return W90GPU_ERR_FILE;
  }
  const std::string read_err = "Error: Problem reading input file " + name;
  long sz[3];
  bool ok = false;
  const char* body = f.n ? read_sizes(f, sz, &ok) : nullptr;
  if (!ok) {
    set_msg(errmsg, errlen, read_err);
    return W90GPU_ERR_FILE;
  }
  if (sz[0] != num_bands) {
    set_msg(errmsg, errlen, name + " has not the right number of bands");
    return W90GPU_ERR_FILE;
  }
  if (sz[1] != num_kpts) {
    set_msg(errmsg, errlen, name + " has not the right number of k-points");
    return W90GPU_ERR_FILE;
  }
  if (sz[2] != num_proj) {
    set_msg(errmsg, errlen, name + " has not the right number of projections");
    return W90GPU_ERR_FILE;
  }
  const long nb = num_bands, np = num_proj, nk = num_kpts;
  const long nrec = nb * np * nk;
  const char* end = f.p + f.n;
  const int nt = pick_threads(nthreads, (size_t)(end - body));
  const std::vector<const char*> cut = cut_ranges(body, end, nt);
  const std::vector<long> cnt = count_records(cut);
  std::vector<long> r0(nt + 1, 0);
  for (int i = 0; i < nt; ++i) r0[i + 1] = r0[i] + cnt[i];
  if (r0[nt] < nrec) {
    set_msg(errmsg, errlen, read_err);
    return W90GPU_ERR_FILE;
  }
  // each record's destination and value, scattered in record order after
  // the parallel parse (the later of two records with one destination wins)
  std::vector<long> dst(nrec);
  std::vector<double> val(2 * (size_t)nrec);
  std::vector<Err> errs(nt);
  run_threads(nt, [&](int i) {
    Err& er = errs[i];
    long r = r0[i];
    for (const char* s = cut[i]; s < cut[i + 1] && r < nrec;) {
      const char* le = line_end(s, cut[i + 1]);
      if (blank_line(s, le)) {
        s = le + 1;
        continue;
      }
      const char* c = s;
      long v[3];
      double x[2];
      bool good = true;
      for (int j = 0; j < 3 && good; ++j) {
        const char *b, *t;
        good = next_token(c, le, &b, &t) && parse_int(b, t, &v[j]);
      }
      for (int j = 0; j < 2 && good; ++j) {
        const char *b, *t;
        good = next_token(c, le, &b, &t) && parse_real(b, t, &x[j]);
      }
      if (good && (v[0] < 1 || v[0] > nb || v[1] < 1 || v[1] > np || v[2] < 1 || v[2] > nk)) good = false;
      if (!good) {
        er.set(r, read_err);
        dst[r] = -1;
      } else {
        dst[r] = (v[0] - 1) + nb * ((v[1] - 1) + np * (v[2] - 1));
        val[2 * r] = x[0];
        val[2 * r + 1] = x[1];
      }
      ++r;
      s = le + 1;
    }
  });
  Err first;
  for (const Err& er : errs)
    if (er.rec >= 0) first.set(er.rec, er.msg);
  if (first.rec >= 0) {
    set_msg(errmsg, errlen, first.msg);
    return W90GPU_ERR_FILE;
  }
  std::memset(a_matrix, 0, sizeof(double) * 2 * (size_t)nrec);
  for (long r = 0; r < nrec; ++r) {
    a_matrix[2 * dst[r]] = val[2 * r];
    a_matrix[2 * dst[r] + 1] = val[2 * r + 1];
  }
  return W90GPU_OK;
}

}  // namespace

// The entry points: no exception leaves the C ABI (an allocation that
// fails is W90GPU_ERR_ALLOC).
extern "C" int w90gpu_read_mmn(const char* path, int num_bands, int num_kpts, int nntot, const int* nnlist,
                               const int* nncell, double* m_matrix, int nthreads, char* errmsg, int errlen) {
  try {
    return read_mmn(path, num_bands, num_kpts, nntot, nnlist, nncell, m_matrix, nthreads, errmsg, errlen);
  } catch (...) {
    set_msg(errmsg, errlen, "w90gpu_read_mmn: out of host memory");
    return W90GPU_ERR_ALLOC;
  }
}

extern "C" int w90gpu_read_amn(const char* path, int num_bands, int num_kpts, int num_proj, double* a_matrix,
                               int nthreads, char* errmsg, int errlen) {
  try {
    return read_amn(path, num_bands, num_kpts, num_proj, a_matrix, nthreads, errmsg, errlen);
  } catch (...) {
    set_msg(errmsg, errlen, "w90gpu_read_amn: out of host memory");
    return W90GPU_ERR_ALLOC;
  }
}
