// deck_io.cpp — host-side text parsing for the reference's input decks: the
// grid file is whitespace-separated decimal coordinates read with
// operator>> (Mesh.cpp:587-607); a 256x128x128 block is 12.6M numbers.  The
// text is cut at whitespace into one chunk per thread; each thread counts its
// tokens, then parses them with std::from_chars (correctly rounded, like
// strtod) into its slice of the output.
#include <algorithm>
#include <cctype>
#include <charconv>
#include <thread>
#include <vector>

#include "../../include/cfx.h"

namespace {

inline bool is_space(char c) { return c == ' ' || c == '\n' || c == '\t' || c == '\r' || c == '\v' || c == '\f'; }

long long count_tokens(const char *a, const char *b) {
  long long n = 0;
  bool in = false;
  for (const char *p = a; p < b; ++p) {
    const bool s = is_space(*p);
    n += (!s && !in);
    in = !s;
  }
  return n;
}

// Parses the tokens of [a, b) into out[0..n_max); returns the count or, for a
// token that is not a whole number, -1 - (its index within the chunk).
long long parse_tokens(const char *a, const char *b, double *out, long long n_max) {
  long long n = 0;
  const char *p = a;
  while (n < n_max) {
    while (p < b && is_space(*p)) ++p;
    if (p >= b) break;
    const char *e = p;
    while (e < b && !is_space(*e)) ++e;
    const char *q = (*p == '+' && e - p > 1) ? p + 1 : p;  // operator>> takes a leading '+'
    double v = 0;
    const auto r = std::from_chars(q, e, v);
    if (r.ec != std::errc() || r.ptr != e) return -1 - n;
    out[n++] = v;
    p = e;
  }
  return n;
}

}  // namespace

int cfx_parse_numbers(const char *text, long long len, double *out, long long n_max, long long *n_out) {
  if (!text || len < 0 || (!out && n_max > 0) || !n_out) return CFX_E_ARG;
  const long long per_thread = 1 << 22;  // 4 MiB of text per thread at least
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  const int T = (int)std::min<long long>(std::min(16u, hw), std::max(1LL, len / per_thread));
  std::vector<const char *> cut(T + 1);
  cut[0] = text;
  cut[T] = text + len;
  for (int t = 1; t < T; ++t) {  // chunk ends moved forward to a whitespace character
    const char *c = std::max(cut[t - 1], text + len * t / T);
    while (c < text + len && !is_space(*c)) ++c;
    cut[t] = c;
  }
  std::vector<long long> cnt(T, 0), res(T, 0);
  auto run = [&](auto &&fn) {
    std::vector<std::thread> th;
    for (int t = 1; t < T; ++t) th.emplace_back(fn, t);
    fn(0);
    for (auto &x : th) x.join();
  };
  run([&](int t) { cnt[t] = count_tokens(cut[t], cut[t + 1]); });
  std::vector<long long> off(T + 1, 0);
  for (int t = 0; t < T; ++t) off[t + 1] = off[t] + cnt[t];
  run([&](int t) {
    const long long lo = std::min(off[t], n_max), hi = std::min(off[t + 1], n_max);
    res[t] = hi > lo ? parse_tokens(cut[t], cut[t + 1], out + lo, hi - lo) : 0;
  });
  for (int t = 0; t < T; ++t)
    if (res[t] < 0) {
      *n_out = off[t] + (-1 - res[t]);  // index of the first malformed token
      return CFX_E_ARG;
    }
  *n_out = std::min(off[T], n_max);
  return 0;
}
