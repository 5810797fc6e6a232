// TZif reader and POSIX TZ rules (see tz.h).
#include "tz.h"

#include <algorithm>
#include <cstdlib>
#include <fstream>
#include <iterator>
#include <sstream>
#include <stdexcept>

namespace nigiri_amd {

// days since 1970-01-01 of y-m-d, and back (the era-based conversion of the
// proleptic Gregorian calendar: 400-year eras of 146097 days)
std::int32_t days_from_civil(int y, unsigned m, unsigned d) {
  y -= m <= 2U ? 1 : 0;
  int const era = (y >= 0 ? y : y - 399) / 400;
  auto const yoe = static_cast<unsigned>(y - era * 400);
  auto const doy = (153U * (m > 2U ? m - 3U : m + 9U) + 2U) / 5U + d - 1U;
  auto const doe = yoe * 365U + yoe / 4U - yoe / 100U + doy;
  return era * 146097 + static_cast<int>(doe) - 719468;
}

void civil_from_days(std::int32_t z, int& y, unsigned& m, unsigned& d) {
  z += 719468;
  int const era = (z >= 0 ? z : z - 146096) / 146097;
  auto const doe = static_cast<unsigned>(z - era * 146097);
  auto const yoe = (doe - doe / 1460U + doe / 36524U - doe / 146096U) / 365U;
  auto const doy = doe - (365U * yoe + yoe / 4U - yoe / 100U);
  auto const mp = (5U * doy + 2U) / 153U;
  d = doy - (153U * mp + 2U) / 5U + 1U;
  m = mp < 10U ? mp + 3U : mp - 9U;
  y = static_cast<int>(yoe) + era * 400 + (m <= 2U ? 1 : 0);
}

namespace {

bool is_leap(int y) { return (y % 4 == 0 && y % 100 != 0) || y % 400 == 0; }

unsigned days_in_month(int y, unsigned m) {
  static constexpr unsigned kDays[] = {31, 28, 31, 30, 31, 30, 31, 31, 30, 31, 30, 31};
  return m == 2U && is_leap(y) ? 29U : kDays[m - 1U];
}

// weekday of a day count, 0 = Sunday (1970-01-01 was a Thursday)
int weekday(std::int32_t z) { return static_cast<int>(((z % 7) + 11) % 7); }

std::int64_t be(std::string const& b, std::size_t pos, int n) {
  if (pos + static_cast<std::size_t>(n) > b.size()) throw std::invalid_argument("TZif: truncated");
  std::uint64_t v = 0U;
  for (int i = 0; i != n; ++i) v = (v << 8) | static_cast<unsigned char>(b[pos + static_cast<std::size_t>(i)]);
  if (n == 4) return static_cast<std::int32_t>(static_cast<std::uint32_t>(v));
  return static_cast<std::int64_t>(v);
}

// [+-]hh[:mm[:ss]] in seconds
std::int32_t parse_hms(std::string const& s, std::size_t& i) {
  int sign = 1;
  if (i < s.size() && (s[i] == '+' || s[i] == '-')) {
    sign = s[i] == '-' ? -1 : 1;
    ++i;
  }
  auto const num = [&]() {
    if (i >= s.size() || s[i] < '0' || s[i] > '9') throw std::invalid_argument("TZ rule: number expected");
    int v = 0;
    while (i < s.size() && s[i] >= '0' && s[i] <= '9') v = v * 10 + (s[i++] - '0');
    return v;
  };
  std::int32_t v = num() * 3600;
  if (i < s.size() && s[i] == ':') {
    ++i;
    v += num() * 60;
    if (i < s.size() && s[i] == ':') {
      ++i;
      v += num();
    }
  }
  return sign * v;
}

void skip_name(std::string const& s, std::size_t& i) {
  if (i < s.size() && s[i] == '<') {
    while (i < s.size() && s[i] != '>') ++i;
    if (i == s.size()) throw std::invalid_argument("TZ rule: unterminated name");
    ++i;
  } else {
    auto const b = i;
    while (i < s.size() && ((s[i] >= 'A' && s[i] <= 'Z') || (s[i] >= 'a' && s[i] <= 'z'))) ++i;
    if (i - b < 3U) throw std::invalid_argument("TZ rule: bad name");
  }
}

posix_rule::date_rule parse_date_rule(std::string const& s, std::size_t& i) {
  posix_rule::date_rule r;
  auto const num = [&]() {
    int v = 0;
    if (i >= s.size() || s[i] < '0' || s[i] > '9') throw std::invalid_argument("TZ rule: bad date");
    while (i < s.size() && s[i] >= '0' && s[i] <= '9') v = v * 10 + (s[i++] - '0');
    return v;
  };
  if (s[i] == 'M') {
    ++i;
    r.kind = 'M';
    r.m = num();
    if (i >= s.size() || s[i++] != '.') throw std::invalid_argument("TZ rule: bad M date");
    r.w = num();
    if (i >= s.size() || s[i++] != '.') throw std::invalid_argument("TZ rule: bad M date");
    r.d = num();
    if (r.m < 1 || r.m > 12 || r.w < 1 || r.w > 5 || r.d > 6) throw std::invalid_argument("TZ rule: bad M date");
  } else if (s[i] == 'J') {
    ++i;
    r.kind = 'J';
    r.n = num();
  } else {
    r.kind = 'D';
    r.n = num();
  }
  r.time = 7200;
  if (i < s.size() && s[i] == '/') {
    ++i;
    r.time = parse_hms(s, i);
  }
  return r;
}

// day count of the rule's date in year y
std::int32_t rule_day(posix_rule::date_rule const& r, int y) {
  auto const jan1 = days_from_civil(y, 1U, 1U);
  switch (r.kind) {
    case 'J': return jan1 + r.n - 1 + (is_leap(y) && r.n >= 60 ? 1 : 0);
    case 'D': return jan1 + r.n;
    default: {
      auto const first = days_from_civil(y, static_cast<unsigned>(r.m), 1U);
      auto day = 1 + (r.d - weekday(first) + 7) % 7 + (r.w - 1) * 7;
      while (day > static_cast<int>(days_in_month(y, static_cast<unsigned>(r.m)))) day -= 7;
      return first + day - 1;
    }
  }
}

std::int32_t rule_offset_at(posix_rule const& p, std::int64_t t) {
  if (!p.has_dst) return p.std_off;
  std::int64_t const local_std = t + p.std_off;
  auto const day = static_cast<std::int32_t>(local_std >= 0 ? local_std / 86400 : -((-local_std + 86399) / 86400));
  int y = 0;
  unsigned m = 0U, d = 0U;
  civil_from_days(day, y, m, d);
  // instants (UTC) of this year's transitions: the start is given in local
  // standard time, the end in local daylight time
  std::int64_t const start = static_cast<std::int64_t>(rule_day(p.start, y)) * 86400 + p.start.time - p.std_off;
  std::int64_t const end = static_cast<std::int64_t>(rule_day(p.end, y)) * 86400 + p.end.time - p.dst_off;
  bool const dst = start < end ? (t >= start && t < end) : !(t >= end && t < start);
  return dst ? p.dst_off : p.std_off;
}

}  // namespace

posix_rule parse_posix_tz(std::string const& s) {
  posix_rule p;
  if (s.empty()) return p;
  std::size_t i = 0U;
  skip_name(s, i);
  p.std_off = -parse_hms(s, i);  // POSIX offsets count west of Greenwich
  p.valid = true;
  if (i == s.size()) return p;
  skip_name(s, i);
  p.has_dst = true;
  p.dst_off = p.std_off + 3600;
  if (i < s.size() && s[i] != ',') p.dst_off = -parse_hms(s, i);
  if (i == s.size()) {  // no rule given: the historical US default
    std::size_t j = 0U;
    std::string const def = "M3.2.0,M11.1.0";
    p.start = parse_date_rule(def, j);
    ++j;
    p.end = parse_date_rule(def, j);
    return p;
  }
  if (s[i++] != ',') throw std::invalid_argument("TZ rule: ',' expected");
  p.start = parse_date_rule(s, i);
  if (i >= s.size() || s[i++] != ',') throw std::invalid_argument("TZ rule: ',' expected");
  p.end = parse_date_rule(s, i);
  return p;
}

time_zone parse_tzif(std::string const& name, std::string const& b) {
  if (b.size() < 44U || b.compare(0, 4, "TZif") != 0) throw std::invalid_argument("TZif: bad magic in " + name);
  auto const version = b[4];
  auto const counts = [&](std::size_t at, std::int64_t c[6]) {
    for (int k = 0; k != 6; ++k) c[k] = be(b, at + 20U + 4U * static_cast<std::size_t>(k), 4);
  };
  std::int64_t c[6];  // isutcnt, isstdcnt, leapcnt, timecnt, typecnt, charcnt
  counts(0U, c);
  auto const block = [&](std::int64_t const cc[6], int tsize) {
    return static_cast<std::size_t>(cc[3] * tsize + cc[3] + cc[4] * 6 + cc[5] + cc[2] * (tsize + 4) + cc[1] + cc[0]);
  };
  time_zone tz;
  tz.name = name;
  std::size_t pos = 44U;
  int tsize = 4;
  if (version >= '2') {  // skip the 32-bit block, use the 64-bit one
    pos += block(c, 4);
    if (pos + 44U > b.size() || b.compare(pos, 4, "TZif") != 0) throw std::invalid_argument("TZif: bad v2 header");
    counts(pos, c);
    pos += 44U;
    tsize = 8;
  }
  auto const timecnt = static_cast<std::size_t>(c[3]), typecnt = static_cast<std::size_t>(c[4]);
  if (typecnt == 0U) throw std::invalid_argument("TZif: no local time types");
  for (std::size_t k = 0; k != timecnt; ++k) tz.transitions.push_back(be(b, pos + k * static_cast<std::size_t>(tsize), tsize));
  pos += timecnt * static_cast<std::size_t>(tsize);
  for (std::size_t k = 0; k != timecnt; ++k) {
    auto const t = static_cast<std::uint8_t>(b.at(pos + k));
    if (t >= typecnt) throw std::invalid_argument("TZif: bad type index");
    tz.trans_type.push_back(t);
  }
  pos += timecnt;
  for (std::size_t k = 0; k != typecnt; ++k) tz.type_off.push_back(static_cast<std::int32_t>(be(b, pos + 6U * k, 4)));
  pos += block(c, tsize) - timecnt * static_cast<std::size_t>(tsize) - timecnt;
  if (version >= '2' && pos < b.size() && b[pos] == '\n') {
    auto const e = b.find('\n', pos + 1U);
    if (e != std::string::npos) tz.footer = parse_posix_tz(b.substr(pos + 1U, e - pos - 1U));
  }
  return tz;
}

time_zone load_time_zone(std::string const& dir, std::string const& name) {
  if (name.empty() || name.find("..") != std::string::npos) throw std::invalid_argument("bad time zone name \"" + name + "\"");
  std::string base = dir;
  if (base.empty()) {
    auto const* env = std::getenv("TZDIR");
    base = env != nullptr && *env != '\0' ? env : "/usr/share/zoneinfo";
  }
  std::ifstream in{base + "/" + name, std::ios::binary};
  if (!in) throw std::invalid_argument("time zone \"" + name + "\" not found in " + base);
  std::string bytes{std::istreambuf_iterator<char>{in}, std::istreambuf_iterator<char>{}};
  return parse_tzif(name, bytes);
}

std::int32_t time_zone::offset_at(std::int64_t t) const {
  if (transitions.empty() || t >= transitions.back()) {
    if (footer.valid) return rule_offset_at(footer, t);
    if (transitions.empty()) return type_off.front();
  }
  if (t < transitions.front()) return type_off.front();
  auto const it = std::upper_bound(transitions.begin(), transitions.end(), t);
  return type_off[trans_type[static_cast<std::size_t>(it - transitions.begin()) - 1U]];
}

std::int32_t noon_offset_minutes(time_zone const& tz, std::int32_t day) {
  // the offset o with offset_at(local noon - o) == o; the candidates are the
  // zone's local time types and its rule's offsets
  std::vector<std::int32_t> cand = tz.type_off;
  if (tz.footer.valid) {
    cand.push_back(tz.footer.std_off);
    if (tz.footer.has_dst) cand.push_back(tz.footer.dst_off);
  }
  std::int64_t const noon = static_cast<std::int64_t>(day) * 86400 + 12 * 3600;
  for (auto const o : cand) {
    if (tz.offset_at(noon - o) == o) return o / 60;
  }
  return tz.offset_at(noon) / 60;
}

}  // namespace nigiri_amd
