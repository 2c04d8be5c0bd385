// Share files in the reference's CSV formats (host code around the device
// vectors).  In the reference each party process reads and writes its own
// file; one call here covers all three parties, so it takes the three
// parties' file paths, party p's file holding its replicated pair
// (s_p, s_{p+1}).
//
//   EVector(size, file) / EVector::output (include/core/containers/e_vector.h:249-348):
//     one row per element, "a,b" (comma or space separated); a first line
//     with other characters than " -0123456789" is skipped; rows past the
//     vector's size are ignored and missing rows stay zero shares.
//   EncodedTable::inputCSVTableSecretShares / outputCSVTableSecretShares
//     (include/core/containers/tabular/encoded_table.h:491-628): a header
//     "[COL]_0,[COL]_1,..." (the replication index after the last '_') and
//     one row per table row; a table whose valid column "[##V]" is not in
//     the file gets it as party 0's secret sharing of ones over the rows read.
//   EncodedTable::inputCSVTableData (encoded_table.h:391-488): a plaintext
//     CSV with a header of column names, read by the input party; every
//     schema column is then secret-shared from that party, and "[##V]", when
//     the file lacks it, is 1 on the rows read.
//
// Values are decimal, signed in the column's share width, as the reference's
// streams print int32_t / int64_t shares.
#include <algorithm>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "internal.h"

namespace orqg {
namespace {

const char *kValid = "[##V]";

std::vector<std::string> split(const std::string &line, char sep) {
  std::vector<std::string> out;
  std::stringstream ss(line);
  std::string tok;
  while (std::getline(ss, tok, sep)) out.push_back(tok);
  return out;
}

// std::stoll as the reference parses a token, then the cast to the share type
uint64_t parse_value(const std::string &tok, const std::string &file) {
  const char *s = tok.c_str();
  char *end = nullptr;
  errno = 0;
  const long long v = std::strtoll(s, &end, 10);
  if (end == s || errno == ERANGE) throw Error(ORQG_E_ARG, "csv: bad value '" + tok + "' in " + file);
  return (uint64_t)v;
}

void put_value(std::ostream &os, uint64_t v, int width) {
  if (width == 4)
    os << (int32_t)(uint32_t)v;
  else
    os << (int64_t)v;
}

// host [3][n] share image of a device vector, and back
std::vector<uint64_t> host_shares(orqg_ctx *c, const orqg_vec *v) {
  const size_t n = v->n, w = v->width;
  std::vector<uint8_t> raw(3 * n * w);
  if (n) {
    ORQG_HIP(hipMemcpyAsync(raw.data(), v->d, raw.size(), hipMemcpyDeviceToHost, c->stream));
    ORQG_HIP(hipStreamSynchronize(c->stream));
  }
  check_swaps(c);
  std::vector<uint64_t> s(3 * n);
  for (size_t i = 0; i < n; i++)
    for (int k = 0; k < 3; k++) {
      if (w == 4) {
        uint32_t x;
        std::memcpy(&x, &raw[(3 * i + k) * 4], 4);
        s[k * n + i] = x;
      } else {
        std::memcpy(&s[k * n + i], &raw[(3 * i + k) * 8], 8);
      }
    }
  return s;
}

void put_shares(orqg_ctx *c, orqg_vec *v, const std::vector<uint64_t> &s) {
  const size_t n = v->n, w = v->width;
  std::vector<uint8_t> raw(3 * n * w);
  for (size_t i = 0; i < n; i++)
    for (int k = 0; k < 3; k++) {
      if (w == 4) {
        const uint32_t x = (uint32_t)s[k * n + i];
        std::memcpy(&raw[(3 * i + k) * 4], &x, 4);
      } else {
        std::memcpy(&raw[(3 * i + k) * 8], &s[k * n + i], 8);
      }
    }
  if (n) {
    ORQG_HIP(hipMemcpyAsync(v->d, raw.data(), raw.size(), hipMemcpyHostToDevice, c->stream));
    ORQG_HIP(hipStreamSynchronize(c->stream));
  }
}

uint64_t trunc_to(uint64_t v, int width) { return width == 4 ? (uint64_t)(uint32_t)v : v; }

// the three parties' pairs -> the three distinct shares; the two holders of
// each share must agree
void merge_pairs(const std::vector<uint64_t> (&pair0)[3], const std::vector<uint64_t> (&pair1)[3],
                 size_t n, std::vector<uint64_t> &s, const std::string &what) {
  s.assign(3 * n, 0);
  for (int p = 0; p < 3; p++)
    for (size_t i = 0; i < n; i++) {
      s[p * n + i] = pair0[p][i];
      if (pair1[(p + 2) % 3][i] != pair0[p][i])
        throw Error(ORQG_E_PARITY, what + ": parties " + std::to_string((p + 2) % 3) + " and " +
                                       std::to_string(p) + " hold different copies of share " +
                                       std::to_string(p) + " in row " + std::to_string(i));
    }
}

}  // namespace

void vec_read_share_files(orqg_ctx *c, orqg_vec *v, const char *const files[3]) {
  const size_t n = v->n;
  std::vector<uint64_t> a[3], b[3];
  for (int p = 0; p < 3; p++) {
    a[p].assign(n, 0);
    b[p].assign(n, 0);
    std::ifstream in(files[p]);
    if (!in.is_open()) throw Error(ORQG_E_ARG, std::string("csv: cannot open ") + files[p]);
    std::string line;
    size_t i = 0;
    while (i < n && std::getline(in, line)) {
      std::replace(line.begin(), line.end(), ',', ' ');
      if (line.find_first_not_of(" -0123456789") != std::string::npos) continue;  // a header
      std::istringstream iss(line);
      std::string t0, t1;
      if (!(iss >> t0 >> t1)) throw Error(ORQG_E_ARG, std::string("csv: short row in ") + files[p]);
      a[p][i] = trunc_to(parse_value(t0, files[p]), v->width);
      b[p][i] = trunc_to(parse_value(t1, files[p]), v->width);
      i++;
    }
  }
  std::vector<uint64_t> s;
  merge_pairs(a, b, n, s, "vec_read_share_files");
  put_shares(c, v, s);
}

void vec_write_share_files(orqg_ctx *c, const orqg_vec *v, const char *const files[3]) {
  const size_t n = v->n;
  const std::vector<uint64_t> s = host_shares(c, v);
  for (int p = 0; p < 3; p++) {
    std::ofstream out(files[p], std::ios::out | std::ios::trunc);
    if (!out.is_open()) throw Error(ORQG_E_ARG, std::string("csv: cannot write ") + files[p]);
    for (size_t i = 0; i < n; i++) {
      put_value(out, s[p * n + i], v->width);
      out << ',';
      put_value(out, s[((p + 1) % 3) * n + i], v->width);
      out << '\n';
    }
  }
}

void table_write_share_csv(orqg_ctx *c, const char *const files[3], const char *const *names,
                           orqg_vec *const *cols, size_t ncols) {
  const size_t n = ncols ? cols[0]->n : 0;
  std::vector<std::vector<uint64_t>> s(ncols);
  for (size_t j = 0; j < ncols; j++) s[j] = host_shares(c, cols[j]);
  for (int p = 0; p < 3; p++) {
    std::ofstream out(files[p], std::ios::out | std::ios::trunc);
    if (!out.is_open()) throw Error(ORQG_E_ARG, std::string("csv: cannot write ") + files[p]);
    for (size_t j = 0; j < ncols; j++) {
      out << names[j] << "_0," << names[j] << "_1";
      out << (j + 1 < ncols ? "," : "\n");
    }
    if (!ncols) out << '\n';
    for (size_t i = 0; i < n; i++)
      for (size_t j = 0; j < ncols; j++) {
        put_value(out, s[j][p * n + i], cols[j]->width);
        out << ',';
        put_value(out, s[j][((p + 1) % 3) * n + i], cols[j]->width);
        out << (j + 1 < ncols ? "," : "\n");
      }
  }
}

void table_read_share_csv(orqg_ctx *c, const char *const files[3], const char *const *names,
                          orqg_vec *const *cols, size_t ncols) {
  if (!ncols) return;
  const size_t n = cols[0]->n;
  std::map<std::string, size_t> index;
  for (size_t j = 0; j < ncols; j++) index[names[j]] = j;
  // [party][column][replica] rows as read
  std::vector<std::vector<std::vector<uint64_t>>> got(
      3, std::vector<std::vector<uint64_t>>(ncols, std::vector<uint64_t>()));
  std::vector<std::vector<bool>> present(3, std::vector<bool>(ncols, false));
  size_t rows_read[3] = {0, 0, 0};
  for (int p = 0; p < 3; p++) {
    for (size_t j = 0; j < ncols; j++) got[p][j].assign(2 * n, 0);
    std::ifstream in(files[p]);
    if (!in.is_open()) throw Error(ORQG_E_ARG, std::string("csv: cannot open ") + files[p]);
    std::string line;
    if (!std::getline(in, line)) throw Error(ORQG_E_ARG, std::string("csv: empty ") + files[p]);
    std::vector<std::pair<long, int>> map;  // per header token: (column, replica) or -1
    for (const std::string &tok : split(line, ',')) {
      std::string name = tok;
      int rep = 0;
      const size_t pos = tok.find_last_of('_');
      if (pos != std::string::npos) {
        name = tok.substr(0, pos);
        rep = std::atoi(tok.substr(pos + 1).c_str());
      }
      auto it = index.find(name);
      if (it == index.end() || rep < 0 || rep > 1)
        throw Error(ORQG_E_ARG, "csv: column " + tok + " of " + files[p] + " is not in the table");
      map.push_back({(long)it->second, rep});
      present[p][it->second] = true;
    }
    size_t i = 0;
    while (i < n && std::getline(in, line)) {
      const std::vector<std::string> toks = split(line, ',');
      for (size_t t = 0; t < toks.size() && t < map.size(); t++) {
        const size_t j = (size_t)map[t].first;
        got[p][j][map[t].second * n + i] = trunc_to(parse_value(toks[t], files[p]), cols[j]->width);
      }
      i++;
    }
    rows_read[p] = i;
  }
  for (size_t j = 0; j < ncols; j++) {
    if (present[0][j] != present[1][j] || present[0][j] != present[2][j])
      throw Error(ORQG_E_ARG, std::string("csv: column ") + names[j] + " is in some parties' files only");
    if (!present[0][j]) continue;
    std::vector<uint64_t> a[3], b[3];
    for (int p = 0; p < 3; p++) {
      a[p].assign(got[p][j].begin(), got[p][j].begin() + n);
      b[p].assign(got[p][j].begin() + n, got[p][j].end());
    }
    std::vector<uint64_t> s;
    merge_pairs(a, b, n, s, std::string("table_read_share_csv column ") + names[j]);
    put_shares(c, cols[j], s);
  }
  // the valid column when the files lack it: party 0 shares ones over the rows read
  auto v = index.find(kValid);
  if (v != index.end() && !present[0][v->second]) {
    const size_t rows = rows_read[0];
    orqg_vec *valid = cols[v->second];
    std::vector<uint8_t> ones(rows * valid->width, 0);
    for (size_t i = 0; i < rows; i++) ones[i * valid->width] = 1;
    orqg_vec head = *valid;  // a view of rows [0, rows)
    head.n = rows;
    if (rows) {
      secret_share(c, ones.data(), &head, 0);
      ORQG_HIP(hipStreamSynchronize(c->stream));  // the copy out of `ones` is done
    }
  }
}

void table_read_csv_data(orqg_ctx *c, const char *path, const char *const *names,
                         orqg_vec *const *cols, size_t ncols, int input_party) {
  if (!ncols) return;
  if (input_party < 0 || input_party > 2) throw Error(ORQG_E_ARG, "csv: input party must be 0..2");
  const size_t n = cols[0]->n;
  std::map<std::string, size_t> index;
  for (size_t j = 0; j < ncols; j++) index[names[j]] = j;
  std::vector<std::vector<uint64_t>> data(ncols, std::vector<uint64_t>(n, 0));
  std::ifstream in(path);
  if (!in.is_open()) throw Error(ORQG_E_ARG, std::string("csv: cannot open ") + path);
  std::string line;
  std::vector<long> col_of;
  std::vector<bool> read(ncols, false);
  if (std::getline(in, line))
    for (const std::string &tok : split(line, ',')) {
      auto it = index.find(tok);
      if (it == index.end()) throw Error(ORQG_E_ARG, "csv: column " + tok + " not found in schema");
      col_of.push_back((long)it->second);
      read[it->second] = true;
    }
  size_t rows = 0;
  while (rows < n && std::getline(in, line)) {
    const std::vector<std::string> toks = split(line, ',');
    for (size_t t = 0; t < toks.size() && t < col_of.size(); t++)
      data[col_of[t]][rows] = trunc_to(parse_value(toks[t], path), cols[col_of[t]]->width);
    rows++;
  }
  auto v = index.find(kValid);
  if (v != index.end() && !read[v->second])
    for (size_t i = 0; i < rows; i++) data[v->second][i] = 1;
  // secret-share every column from the input party, in the order given
  for (size_t j = 0; j < ncols; j++) {
    const int w = cols[j]->width;
    std::vector<uint8_t> plain(n * w);
    for (size_t i = 0; i < n; i++) std::memcpy(&plain[i * w], &data[j][i], w);
    secret_share(c, plain.data(), cols[j], input_party);
    ORQG_HIP(hipStreamSynchronize(c->stream));  // the copy out of `plain` is done
  }
}

}  // namespace orqg
