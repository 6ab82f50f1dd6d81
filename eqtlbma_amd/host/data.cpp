#include "data.hpp"

#include <sys/stat.h>
#include <zlib.h>

#include <algorithm>
#include <climits>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <unordered_map>

#include "util.hpp"

namespace host {
namespace {

const double kNaN = std::numeric_limits<double>::quiet_NaN();

bool is_missing(const std::string& t) { return t == "NA" || t == "na" || t == "NaN" || t == "nan"; }
bool is_id(const std::string& t) { return t == "Id" || t == "id" || t == "ID"; }

/* subgroup -> path list (data_loader.cpp:37-79): exactly two tokens per
 * line (delimiters space, tab, comma), '#' comments, first entry wins. */
std::map<std::string, std::string> load_list(const std::string& file, int verbose) {
  std::map<std::string, std::string> m;
  if (file.empty()) return m;
  LineReader r(file);
  if (verbose > 0) std::cout << "load file " << file << " ..." << std::endl;
  std::string line;
  std::vector<std::string> tok;
  size_t nl = 0;
  while (r.next(line)) {
    ++nl;
    tokenize(line, " \t,", tok);
    if (tok.size() != 2) fail("file " + file + " should have only two columns at line " + std::to_string(nl));
    if (tok[0][0] == '#') continue;
    m.emplace(tok[0], tok[1]);
  }
  if (verbose > 0) std::cout << "items loaded: " << m.size() << std::endl;
  return m;
}

std::vector<std::string> header_samples(const std::string& line, const std::string& file) {
  std::vector<std::string> tok;
  tokenize(line, " \t", tok);
  if (!tok.empty() && is_id(tok[0])) tok.erase(tok.begin());
  std::vector<std::string> s = tok;
  std::sort(s.begin(), s.end());
  if (std::adjacent_find(s.begin(), s.end()) != s.end()) fail("file " + file + " has redundant samples in its header");
  return tok;
}

std::vector<std::string> first_line_samples(const std::string& file) {
  LineReader r(file);
  std::string line;
  if (!r.next(line)) fail("problem with the header of file " + file);
  if (line.empty()) fail("file " + file + " is empty");
  return header_samples(line, file);
}

enum GenoFormat { FMT_CUSTOM, FMT_VCF, FMT_IMPUTE };

bool impute_header(const std::vector<std::string>& t) {
  return t.size() >= 5 && t[0] == "chr" && (t[1] == "name" || t[1] == "id") && t[2] == "coord" && t[3] == "a1" &&
         t[4] == "a2";
}

/* Samples of a genotype file (data_loader.cpp:232-337). */
std::vector<std::string> geno_samples(const std::string& file) {
  LineReader r(file);
  std::string line;
  std::vector<std::string> tok;
  if (!r.next(line)) fail("problem with the header of file " + file);
  if (line.empty()) fail("file " + file + " is empty");
  if (line.find("##fileformat=VCF") != std::string::npos) {
    while (r.next(line)) {
      if (line.find("#CHROM") == std::string::npos) continue;
      tokenize(line, " \t", tok);
      std::vector<std::string> s = tok;
      std::sort(s.begin(), s.end());
      if (std::adjacent_find(s.begin(), s.end()) != s.end()) fail("file " + file + " has redundant samples in its header");
      return std::vector<std::string>(tok.begin() + std::min<size_t>(9, tok.size()), tok.end());
    }
    return {};
  }
  tokenize(line, " \t", tok);
  if (impute_header(tok)) {
    if ((tok.size() - 5) % 3 != 0) fail("the header of IMPUTE file " + file + " is badly formatted");
    std::vector<std::string> out, part;
    for (size_t i = 5; i < tok.size(); i += 3) {
      /* the reference cuts each column name at its first '_' or 'a'
       * (split(tokens[i], "_a", 0), data_loader.cpp:286) */
      tokenize(tok[i], "_a", part);
      if (part.empty()) fail("not enough tokens after splitting string");
      out.push_back(part[0]);
    }
    std::vector<std::string> s = out;
    std::sort(s.begin(), s.end());
    if (std::adjacent_find(s.begin(), s.end()) != s.end()) fail("file " + file + " has redundant samples in its header");
    return out;
  }
  return header_samples(line, file);
}

std::vector<int32_t> columns_of(const std::vector<std::string>& all, const std::vector<std::string>& cols) {
  std::map<std::string, int32_t> pos;
  for (size_t k = 0; k < cols.size(); ++k) pos.emplace(cols[k], (int32_t)k);
  std::vector<int32_t> out(all.size(), -1);
  for (size_t i = 0; i < all.size(); ++i) {
    auto it = pos.find(all[i]);
    if (it != pos.end()) out[i] = it->second;
  }
  return out;
}

double fold_maf(double sum, size_t n) {
  double f = sum / (2.0 * (double)n);
  return f <= 0.5 ? f : 1 - f;
}

/* Adds the genotype row of one SNP to set k; returns its MAF (NaN when a
 * value is missing), as Snp::AddSubgroupFrom*Line (snp.cpp:88-186). */
double add_row(GenoSet& gs, const std::vector<std::string>& tok, size_t first, GenoFormat fmt, size_t idx_gt) {
  const size_t n = (size_t)gs.ncols;
  const size_t base = gs.vals.size();
  gs.vals.resize(base + n, kNaN);
  double sum = 0;
  bool missing = false;
  std::vector<std::string> f2, f3;
  for (size_t i = 0; i < n; ++i) {
    double g = kNaN;
    if (fmt == FMT_CUSTOM) {
      const std::string& t = tok[first + i];
      if (is_missing(t))
        missing = true;
      else
        g = std::atof(t.c_str());
    } else if (fmt == FMT_IMPUTE) {
      const double AA = std::atof(tok[first + 3 * i].c_str()), AB = std::atof(tok[first + 3 * i + 1].c_str()),
                   BB = std::atof(tok[first + 3 * i + 2].c_str());
      (void)AA;
      if (AA == 0 && AB == 0 && BB == 0)
        missing = true;
      else
        g = AB + 2 * BB;
    } else {
      tokenize(tok[first + i], ":", f2);
      if (idx_gt >= f2.size() || f2[idx_gt].find('.') != std::string::npos) {
        missing = true;
      } else {
        tokenize(f2[idx_gt], "|/", f3);
        g = 0;
        if (f3.size() > 0 && f3[0] == "1") g += 1;
        if (f3.size() > 1 && f3[1] == "1") g += 1;
      }
    }
    gs.vals[base + i] = g;
    if (!std::isnan(g)) sum += g;
  }
  return missing ? kNaN : fold_maf(sum, n);
}

void load_snp_coords(const std::string& file, const std::set<std::string>& keep, Data& d) {
  LineReader r(file);
  std::string line;
  std::vector<std::string> tok;
  while (r.next(line)) {
    tokenize(line, " \t", tok);
    if (tok.size() < 4) fail("file " + file + " is not in the BED format");
    if (!keep.empty() && !keep.count(tok[3])) continue;
    if (d.snps.count(tok[3])) continue; /* in case of redundancy */
    if (tok[1] == tok[2]) fail("start and end coordinates of " + tok[3] + " should be different (at least 1 bp)");
    Snp s;
    s.name = tok[3];
    s.chr = tok[0];
    s.pos = std::atol(tok[2].c_str());
    d.snps.emplace(s.name, s);
  }
}

/* ---- tabix-indexed BED (loadSnpInfo with a .tbi, data_loader.cpp:1345-1400)
 *
 * The reference queries the index once per gene (genes in name order) for
 * the region Gene::GetRegionInTabixFormat gives (gene.cpp:94-104:
 * chr:start-r-start+r, or start-r-end+r for TSS+TES, in size_t arithmetic,
 * so start < r wraps), parsed as ti_parse_region does (atoi of each bound,
 * begin - 1 when positive, a region with begin > end dropped, begin clamped
 * at 0 by the query), and keeps every record overlapping it (half-open
 * intervals from the index's column configuration, tabix/index.c:128-208,
 * 860-874); a SNP name seen before is skipped.  Here one pass over the
 * BGZF file gives the same set: a record is kept if it overlaps the region
 * of some gene, and of records sharing a name the one returned first by the
 * per-gene queries -- smallest gene index, then file order -- is kept.  The
 * index file is read for its header only (sequence names, columns,
 * meta character, skipped lines). */
struct TbiHeader {
  int32_t preset = 0, sc = 1, bc = 2, ec = 3, meta = '#', skip = 0;
  std::unordered_map<std::string, int> tid;
};

bool read_tbi(const std::string& path, TbiHeader& h) {
  gzFile f = gzopen(path.c_str(), "rb");
  if (!f) return false;
  char magic[4];
  int32_t v[8];
  bool ok = gzread(f, magic, 4) == 4 && std::memcmp(magic, "TBI\1", 4) == 0 &&
            gzread(f, v, sizeof v) == (int)sizeof v;
  if (ok) {
    h.preset = v[1];
    h.sc = v[2];
    h.bc = v[3];
    h.ec = v[4];
    h.meta = v[5];
    h.skip = v[6];
    const int32_t l_nm = v[7];
    std::vector<char> nm((size_t)std::max(l_nm, 0));
    ok = l_nm >= 0 && gzread(f, nm.data(), (unsigned)l_nm) == l_nm;
    for (int32_t i = 0, start = 0, k = 0; ok && i < l_nm; ++i)
      if (nm[i] == 0) {
        h.tid.emplace(std::string(nm.data() + start, nm.data() + i), k++);
        start = i + 1;
      }
  }
  gzclose(f);
  return ok;
}

struct TbxRegion {
  int beg, end, gene;
};

void load_snp_coords_tabix(const std::string& file, const std::string& idx, const std::set<std::string>& keep,
                           const Data& d, const std::string& anchor, uint64_t radius, Data& out) {
  struct stat st_bed, st_tbi;
  stat(file.c_str(), &st_bed);
  stat(idx.c_str(), &st_tbi);
  if (st_bed.st_mtime > st_tbi.st_mtime) fail("index file (" + idx + ") is older than data file (" + file + ")");
  TbiHeader h;
  if (!read_tbi(idx, h)) fail("failed to load the index file (tabix)");
  if ((h.preset & 0xffff) != 0) fail("failed to load the index file (tabix): not a BED (generic) index");
  /* regions, per sequence of the index */
  std::vector<std::vector<TbxRegion>> reg(h.tid.size());
  int gi = 0;
  for (const auto& kv : d.genes) {
    const Gene& g = kv.second;
    const int gene = gi++;
    auto t = h.tid.find(g.chr);
    if (t == h.tid.end()) continue;
    const uint64_t a = (uint64_t)g.start - radius;
    const uint64_t b = (uint64_t)(anchor == "TSS+TES" ? g.end : g.start) + radius;
    int beg = (int)std::strtol(std::to_string(a).c_str(), nullptr, 10);
    int end = (int)std::strtol(std::to_string(b).c_str(), nullptr, 10);
    if (anchor != "TSS" && anchor != "TSS+TES") {
      beg = 0;
      end = 1 << 29;
    }
    if (beg > 0) --beg;
    if (beg > end) continue;
    if (beg < 0) beg = 0;
    if (end < beg) continue;
    reg[t->second].push_back({beg, end, gene});
  }
  std::vector<std::vector<int>> pmax(reg.size());
  for (size_t t = 0; t < reg.size(); ++t) {
    std::sort(reg[t].begin(), reg[t].end(), [](const TbxRegion& x, const TbxRegion& y) { return x.beg < y.beg; });
    pmax[t].resize(reg[t].size());
    int m = INT_MIN;
    for (size_t i = 0; i < reg[t].size(); ++i) pmax[t][i] = m = std::max(m, reg[t][i].end);
  }
  struct Cand {
    int gene;
    int64_t line;
    std::string chr, coord;
  };
  std::unordered_map<std::string, Cand> best;
  gzFile f = gzopen(file.c_str(), "rb");
  if (!f) fail("fail to open the data file (tabix)");
  gzbuffer(f, 1 << 20);
  std::vector<char> buf(1 << 16);
  std::string line;
  std::vector<std::string> fld, tok;
  int64_t nl = 0;
  while (true) {
    line.clear();
    bool got = false;
    while (gzgets(f, buf.data(), (int)buf.size())) {
      got = true;
      const size_t n = std::strlen(buf.data());
      if (n > 0 && buf[n - 1] == '\n') {
        line.append(buf.data(), n - 1);
        break;
      }
      line.append(buf.data(), n);
    }
    if (!got) break;
    ++nl;
    if (nl <= h.skip || (!line.empty() && (unsigned char)line[0] == (unsigned char)h.meta)) continue;
    /* the record's interval as ti_get_intv forms it (fields split at each
     * tab, empty ones included) */
    fld.clear();
    size_t p0 = 0;
    while (true) {
      const size_t q = line.find('\t', p0);
      fld.emplace_back(line, p0, q == std::string::npos ? std::string::npos : q - p0);
      if (q == std::string::npos) break;
      p0 = q + 1;
    }
    if ((int)fld.size() < h.sc || (int)fld.size() < h.bc) continue;
    auto t = h.tid.find(fld[h.sc - 1]);
    if (t == h.tid.end()) continue;
    long beg = std::strtol(fld[h.bc - 1].c_str(), nullptr, 0), end = beg;
    if (!(h.preset & 0x10000))
      --beg;
    else
      ++end;
    if (beg < 0) beg = 0;
    if (end < 1) end = 1;
    if (h.ec > 0 && h.ec != h.bc && h.ec <= (int)fld.size()) end = std::strtol(fld[h.ec - 1].c_str(), nullptr, 0);
    /* smallest gene index among the regions overlapping [beg, end) */
    const std::vector<TbxRegion>& R = reg[t->second];
    const std::vector<int>& M = pmax[t->second];
    size_t ub = std::lower_bound(R.begin(), R.end(), end, [](const TbxRegion& x, long v) { return x.beg < v; }) -
                R.begin();
    int gmin = INT_MAX;
    for (size_t i = ub; i-- > 0;) {
      if (M[i] <= beg) break;
      if (R[i].end > beg && R[i].gene < gmin) gmin = R[i].gene;
    }
    if (gmin == INT_MAX) continue;
    tokenize(line, "\t", tok);
    if (tok.size() < 4) fail("file " + file + " is not in the BED format");
    if (!keep.empty() && !keep.count(tok[3])) continue;
    auto it = best.find(tok[3]);
    if (it == best.end() || gmin < it->second.gene) best[tok[3]] = Cand{gmin, nl, tok[0], tok[2]};
  }
  gzclose(f);
  for (const auto& kv : best) {
    Snp sn;
    sn.name = kv.first;
    sn.chr = kv.second.chr;
    sn.pos = std::atol(kv.second.coord.c_str());
    out.snps.emplace(sn.name, sn);
  }
}

}  // namespace

bool load_data(const LoadOptions& o, Data& d) {
  const int verbose = o.verbose;
  /* ---- list files and subgroups (data_loader.cpp:81-163) ---- */
  auto exp_files = load_list(o.exp_list, verbose);
  if (!o.subgroups_to_keep.empty())
    for (auto it = exp_files.begin(); it != exp_files.end();)
      it = std::find(o.subgroups_to_keep.begin(), o.subgroups_to_keep.end(), it->first) == o.subgroups_to_keep.end()
               ? exp_files.erase(it)
               : std::next(it);
  auto geno_files = load_list(o.geno_list, verbose);
  for (auto it = geno_files.begin(); it != geno_files.end();) it = exp_files.count(it->first) ? std::next(it) : geno_files.erase(it);
  for (auto it = exp_files.begin(); it != exp_files.end();) it = geno_files.count(it->first) ? std::next(it) : exp_files.erase(it);
  if (o.error_model != "uvlr")
    for (auto& kv : geno_files)
      if (kv.second != geno_files.begin()->second)
        fail("--error mvlr/hybrid requires the same genotypes in a single file for all subgroups");
  auto covar_files = load_list(o.covar_list, verbose);
  for (auto it = covar_files.begin(); it != covar_files.end();) it = exp_files.count(it->first) ? std::next(it) : covar_files.erase(it);
  for (auto& kv : exp_files) {
    Subgroup s;
    s.name = kv.first;
    s.exp_file = kv.second;
    s.geno_file = geno_files[kv.first];
    auto c = covar_files.find(kv.first);
    if (c != covar_files.end()) s.covar_file = c->second;
    d.sbgrps.push_back(s);
  }
  const size_t S = d.sbgrps.size();
  if (verbose > 0) {
    std::cout << "analyze " << S << " subgroup" << (S > 1 ? "s" : "") << " (identifier):" << std::endl;
    for (size_t s = 0; s < S; ++s) std::cout << d.sbgrps[s].name << " (" << s + 1 << ")" << std::endl;
  }
  if (S == 0) return false;

  /* ---- samples (data_loader.cpp:339-394, samples.cpp:60-146) ---- */
  std::vector<std::vector<std::string>> exp_cols(S), geno_cols(S), covar_cols(S);
  std::set<std::string> uni;
  for (size_t s = 0; s < S; ++s) {
    exp_cols[s] = first_line_samples(d.sbgrps[s].exp_file);
    uni.insert(exp_cols[s].begin(), exp_cols[s].end());
  }
  for (size_t s = 0; s < S; ++s) {
    geno_cols[s] = geno_samples(d.sbgrps[s].geno_file);
    uni.insert(geno_cols[s].begin(), geno_cols[s].end());
  }
  d.samples.assign(uni.begin(), uni.end());
  for (size_t s = 0; s < S; ++s)
    if (!d.sbgrps[s].covar_file.empty()) {
      covar_cols[s] = first_line_samples(d.sbgrps[s].covar_file);
      for (auto& n : covar_cols[s])
        if (!uni.count(n)) fail("sample " + n + " has covariates but neither expression levels nor genotypes");
    }
  for (size_t s = 0; s < S; ++s) {
    d.sbgrps[s].exp_col = columns_of(d.samples, exp_cols[s]);
    d.sbgrps[s].geno_col = columns_of(d.samples, geno_cols[s]);
    d.sbgrps[s].covar_col = columns_of(d.samples, covar_cols[s]);
    d.sbgrps[s].exp_ncols = (int32_t)exp_cols[s].size();
  }
  if (verbose > 0) std::cout << "total nb of samples: " << d.samples.size() << std::endl;

  /* ---- covariates (data_loader.cpp:1097-1159, covariates.cpp:56-77) ---- */
  for (size_t s = 0; s < S; ++s) {
    Subgroup& sg = d.sbgrps[s];
    if (sg.covar_file.empty()) continue;
    LineReader r(sg.covar_file);
    std::string line;
    std::vector<std::string> tok;
    r.next(line);
    tokenize(line, " \t", tok);
    const size_t ns = (!tok.empty() && is_id(tok[0])) ? tok.size() - 1 : tok.size();
    std::map<std::string, std::vector<double>> cov;
    size_t nl = 1;
    while (r.next(line)) {
      ++nl;
      tokenize(line, " \t", tok);
      if (tok.size() != ns + 1)
        fail("not enough columns on line " + std::to_string(nl) + " of file " + sg.covar_file + " (" +
             std::to_string(tok.size()) + " != " + std::to_string(ns + 1) + ")");
      if (cov.count(tok[0])) continue;
      std::vector<double> v(ns);
      for (size_t k = 0; k < ns; ++k) {
        if (is_missing(tok[k + 1])) fail("no missing value allowed, see covariate " + tok[0] + " in subgroup " + sg.name);
        v[k] = std::atof(tok[k + 1].c_str());
      }
      cov.emplace(tok[0], v);
    }
    for (auto& kv : cov) {
      sg.covar_names.push_back(kv.first);
      sg.covar_vals.push_back(kv.second);
    }
    if (verbose > 0) std::cout << sg.name << " (" << sg.covar_file << "): " << cov.size() << " covariates" << std::endl;
  }

  /* ---- genes (data_loader.cpp:396-435) ---- */
  std::set<std::string> gene_chrs;
  {
    LineReader r(o.gene_coords);
    std::string line;
    std::vector<std::string> tok;
    while (r.next(line)) {
      tokenize(line, " \t", tok);
      if (tok.size() < 4) fail("file " + o.gene_coords + " is not in the BED format");
      if (d.genes.count(tok[3])) continue;
      if (tok[1] == tok[2]) fail("start and end coordinates of " + tok[3] + " should be different (at least 1 bp)");
      Gene g;
      g.name = tok[3];
      g.chr = tok[0];
      g.start = std::atol(tok[1].c_str()) + 1;
      g.end = std::atol(tok[2].c_str());
      g.exp_row.assign(S, -1);
      gene_chrs.insert(g.chr);
      d.genes.emplace(g.name, g);
    }
    if (verbose > 0) std::cout << "total nb of genes with coordinates: " << d.genes.size() << std::endl;
  }

  /* ---- expression levels (data_loader.cpp:437-527, gene.cpp:104-118) ---- */
  for (size_t s = 0; s < S; ++s) {
    Subgroup& sg = d.sbgrps[s];
    LineReader r(sg.exp_file);
    std::string line;
    std::vector<std::string> tok;
    if (!r.next(line)) fail("problem with the header of file " + sg.exp_file);
    tokenize(line, " \t", tok);
    const size_t ns = (!tok.empty() && is_id(tok[0])) ? tok.size() - 1 : tok.size();
    size_t nl = 1, kept = 0;
    int32_t nrow = 0;
    while (r.next(line)) {
      ++nl;
      tokenize(line, " \t", tok);
      if (tok.size() != ns + 1)
        fail("not enough columns on line " + std::to_string(nl) + " of file " + sg.exp_file + " (" +
             std::to_string(tok.size()) + " != " + std::to_string(ns + 1) + ")");
      auto it = d.genes.find(tok[0]);
      if (it == d.genes.end() || it->second.exp_row[s] >= 0) continue;
      it->second.exp_row[s] = nrow++;
      for (size_t k = 0; k < ns; ++k)
        sg.exp_vals.push_back(is_missing(tok[k + 1]) ? kNaN : std::atof(tok[k + 1].c_str()));
      ++kept;
    }
    sg.exp_ncols = (int32_t)ns;
    if (verbose > 0)
      std::cout << sg.name << " (" << sg.exp_file << "): " << (nl - 1) << " genes (to keep: " << kept << ")"
                << std::endl;
  }
  for (auto it = d.genes.begin(); it != d.genes.end();) {
    bool any = false;
    for (size_t s = 0; s < S; ++s) any |= it->second.exp_row[s] >= 0 && d.sbgrps[s].exp_ncols > 0;
    it = any ? std::next(it) : d.genes.erase(it);
  }
  if (verbose > 0) std::cout << "total nb of genes to analyze: " << d.genes.size() << std::endl;
  if (d.genes.empty()) return false;

  /* ---- SNPs and genotypes (data_loader.cpp:570-1010) ---- */
  std::set<std::string> keep;
  if (!o.snps_to_keep.empty()) {
    LineReader r(o.snps_to_keep);
    std::string line;
    std::vector<std::string> tok;
    size_t nl = 0;
    while (r.next(line)) {
      ++nl;
      tokenize(line, " \t,", tok);
      if (tok.size() != 1) fail("file " + o.snps_to_keep + " should have only one column at line " + std::to_string(nl));
      if (tok[0][0] == '#') continue;
      keep.insert(tok[0]);
    }
    if (verbose > 0) std::cout << "nb of SNPs to keep: " << keep.size() << std::endl;
    if (keep.empty()) return false;
  }
  const bool custom = !o.snp_coords.empty();
  if (custom) {
    if (file_exists(o.snp_coords + ".tbi")) {
      if (verbose > 0) std::cout << "load SNP coordinates (tabix-indexed BED file) ..." << std::endl;
      load_snp_coords_tabix(o.snp_coords, o.snp_coords + ".tbi", keep, d, o.anchor, o.radius, d);
    } else {
      load_snp_coords(o.snp_coords, keep, d);
    }
  }
  /* files are read in subgroup order until the first subgroup (after the
   * first) naming the first subgroup's file; every subgroup from there on
   * takes the first subgroup's genotypes (data_loader.cpp:733-738, 691-715) */
  size_t n_loaded = S;
  for (size_t s = 1; s < S; ++s)
    if (d.sbgrps[s].geno_file == d.sbgrps[0].geno_file) {
      n_loaded = s;
      break;
    }
  for (size_t s = 0; s < S; ++s) d.sbgrps[s].geno_set = s < n_loaded ? (int)s : 0;
  d.sets.resize(n_loaded);
  for (auto& kv : d.snps) {
    kv.second.row.assign(n_loaded, -1);
    kv.second.maf.assign(n_loaded, kNaN);
  }
  for (size_t k = 0; k < n_loaded; ++k) {
    GenoSet& gs = d.sets[k];
    gs.file = d.sbgrps[k].geno_file;
    LineReader r(gs.file);
    std::string line;
    std::vector<std::string> tok, f2;
    if (!r.next(line)) fail("problem with the header of file " + gs.file);
    size_t nl = 1;
    int32_t nrow = 0;
    auto add = [&](Snp& sn, size_t first, GenoFormat fmt, size_t idx_gt) {
      if (sn.row.size() < n_loaded) {
        sn.row.assign(n_loaded, -1);
        sn.maf.assign(n_loaded, kNaN);
      }
      if (sn.row[k] >= 0) fail("SNP " + sn.name + " is duplicated in file " + gs.file);
      sn.maf[k] = add_row(gs, tok, first, fmt, idx_gt);
      sn.row[k] = nrow++;
    };
    if (custom) {
      tokenize(line, " \t", tok);
      if (!tok.empty() && tok[0] == "chr") fail("don't use --scoord if genotypes in IMPUTE format");
      if (!tok.empty() && tok[0].find("##fileformat=VCF") != std::string::npos)
        fail("don't use --scoord if genotypes in VCF format");
      const size_t ns = (!tok.empty() && is_id(tok[0])) ? tok.size() - 1 : tok.size();
      gs.ncols = (int32_t)ns;
      while (r.next(line)) {
        ++nl;
        tokenize(line, " \t", tok);
        if (tok.size() != ns + 1)
          fail("not enough columns on line " + std::to_string(nl) + " of file " + gs.file + " (" +
               std::to_string(tok.size()) + " != " + std::to_string(ns + 1) + ")");
        auto it = d.snps.find(tok[0]);
        if (it == d.snps.end()) continue; /* no coordinate */
        if (std::find(tok.begin(), tok.end(), "NA") != tok.end()) continue;
        add(it->second, 1, FMT_CUSTOM, 0);
      }
    } else if (line.find("##fileformat=VCF") != std::string::npos) {
      while (r.next(line)) {
        ++nl;
        if (line.find("#CHROM") != std::string::npos) break;
      }
      tokenize(line, " \t", tok);
      const size_t ns = tok.size() >= 9 ? tok.size() - 9 : 0;
      gs.ncols = (int32_t)ns;
      while (r.next(line)) {
        ++nl;
        tokenize(line, " \t", tok);
        if (tok.size() != ns + 9)
          fail("not enough columns on line " + std::to_string(nl) + " of file " + gs.file + " (" +
               std::to_string(tok.size()) + " != " + std::to_string(ns + 9) + ")");
        if (tok[8].find("GT") == std::string::npos)
          fail("missing GT in 9-th field on line " + std::to_string(nl) + " of file " + gs.file);
        if (!gene_chrs.count(tok[0])) continue;
        if (!keep.empty() && !keep.count(tok[2])) continue;
        auto it = d.snps.find(tok[2]);
        if (it == d.snps.end()) {
          Snp sn;
          sn.name = tok[2];
          sn.chr = tok[0];
          sn.pos = std::atol(tok[1].c_str());
          it = d.snps.emplace(sn.name, sn).first;
        }
        tokenize(tok[8], ":", f2);
        size_t idx_gt = 0;
        while (idx_gt < f2.size() && f2[idx_gt] != "GT") ++idx_gt;
        add(it->second, 9, FMT_VCF, idx_gt);
      }
    } else {
      tokenize(line, " \t", tok);
      if (impute_header(tok)) {
        if ((tok.size() - 5) % 3 != 0) fail("wrong number of columns on line 1 of file " + gs.file);
        const size_t ns = (tok.size() - 5) / 3;
        gs.ncols = (int32_t)ns;
        while (r.next(line)) {
          ++nl;
          tokenize(line, " \t", tok);
          if (tok.size() != 3 * ns + 5)
            fail("not enough columns on line " + std::to_string(nl) + " of file " + gs.file + " (" +
                 std::to_string(tok.size()) + " != " + std::to_string(3 * ns + 5) + ")");
          if (!gene_chrs.count(tok[0])) continue;
          if (!keep.empty() && !keep.count(tok[1])) continue;
          auto it = d.snps.find(tok[1]);
          if (it == d.snps.end()) {
            Snp sn;
            sn.name = tok[1];
            sn.chr = tok[0];
            sn.pos = std::atol(tok[2].c_str());
            it = d.snps.emplace(sn.name, sn).first;
          }
          add(it->second, 5, FMT_IMPUTE, 0);
        }
      } else if (line.compare(0, 2, "id") == 0 || line.compare(0, 2, "Id") == 0 || line.compare(0, 2, "ID") == 0) {
        fail("file " + gs.file + " seems to be in the custom format but --scoord is missing");
      } else {
        fail("can't recognize the format of file " + gs.file);
      }
    }
    if (verbose > 0) std::cout << d.sbgrps[k].name << " (" << gs.file << "): " << nrow << " SNPs kept" << std::endl;
  }
  /* missing values, then the MAF filter, per genotype set (snp.cpp:232-253) */
  for (auto it = d.snps.begin(); it != d.snps.end();) {
    Snp& sn = it->second;
    bool any = false;
    for (size_t k = 0; k < sn.row.size(); ++k) {
      if (sn.row[k] >= 0 && (std::isnan(sn.maf[k]) || sn.maf[k] < o.min_maf)) sn.row[k] = -1;
      any |= sn.row[k] >= 0 && d.sets[k].ncols > 0;
    }
    it = any ? std::next(it) : d.snps.erase(it);
  }
  for (auto& kv : d.snps) d.chr_snps[kv.second.chr].push_back(&kv.second);
  for (auto& kv : d.chr_snps)
    std::stable_sort(kv.second.begin(), kv.second.end(), [](const Snp* a, const Snp* b) { return a->pos < b->pos; });
  if (verbose > 0) std::cout << "total nb of SNPs to analyze: " << d.snps.size() << std::endl;
  return !d.snps.empty();
}


/* loadListSstatsFile + fillGeneSnpPairsWithSstats (data_loader.cpp:1212-1327):
 * one file per subgroup (subgroups in name order), header columns found by
 * name among tab-separated tokens, genes and pairs created in order of first
 * appearance, the first line of a (gene, SNP, subgroup) kept
 * (GeneSnpPair::SetSstats inserts into maps, gene_snp_pair.cpp:241-254). */
bool load_inss(const std::string& list, int verbose, InssData& d) {
  std::map<std::string, std::string> files;
  {
    LineReader r(list);
    if (verbose > 0) std::cout << "load file " << list << " ..." << std::endl;
    std::string line;
    std::vector<std::string> tok;
    size_t nl = 0;
    while (r.next(line)) {
      ++nl;
      tokenize(line, " \t,", tok);
      if (tok.size() != 2) fail("file " + list + " should have only two columns at line " + std::to_string(nl));
      if (tok[0][0] == '#') continue;
      files.emplace(tok[0], tok[1]);
    }
    if (!r.complete()) fail("can't read successfully file " + list + " up to the end");
    if (verbose > 0) std::cout << "items loaded: " << files.size() << " files" << std::endl;
  }
  d.sbgrps.clear();
  for (auto& kv : files) d.sbgrps.push_back(kv.first);
  const size_t S = d.sbgrps.size();
  static const char* kCols[6] = {"gene", "snp", "n", "sigmahat", "betahat.geno", "sebetahat.geno"};
  std::set<std::string> snp_names;
  size_t s = 0;
  for (auto& kv : files) {
    if (verbose > 0) std::cout << "load summary statistics for subgroup " << kv.first << " ..." << std::endl;
    LineReader r(kv.second);
    std::string line;
    std::vector<std::string> tok;
    size_t col[6];
    size_t need = 0;
    if (!r.next(line)) line.clear();
    tokenize(line, "\t", tok);
    for (int k = 0; k < 6; ++k) {
      col[k] = std::string::npos;
      for (size_t c = 0; c < tok.size(); ++c)
        if (tok[c] == kCols[k]) col[k] = c;
      /* the reference prints the missing index, npos, rather than the name */
      if (col[k] == std::string::npos)
        fail("missing " + std::to_string((unsigned long long)std::string::npos) + " in header of " + kv.second);
      need = std::max(need, col[k] + 1);
    }
    size_t nl = 1;
    while (r.next(line)) {
      ++nl;
      tokenize(line, "\t", tok);
      if (tok.size() < need)
        fail("line " + std::to_string(nl) + " of " + kv.second + " has fewer than " + std::to_string(need) +
             " columns");
      InssGene& g = d.genes[tok[col[0]]];
      const std::string& snp = tok[col[1]];
      auto it = g.idx.find(snp);
      if (it == g.idx.end()) {
        it = g.idx.emplace(snp, g.snps.size()).first;
        g.snps.push_back(snp);
        g.st.emplace_back(S);
        snp_names.insert(snp);
      }
      SstatsEntry& e = g.st[it->second][s];
      if (e.set) continue;
      e.set = true;
      e.n = std::atol(tok[col[2]].c_str());
      e.sigmahat = std::atof(tok[col[3]].c_str());
      e.betahat = std::atof(tok[col[4]].c_str());
      e.sebetahat = std::atof(tok[col[5]].c_str());
    }
    ++s;
  }
  d.n_snps = snp_names.size();
  return !d.genes.empty() && d.n_snps > 0;
}

}  // namespace host
