// Input formats on the host, in C++: FASTA alignment and Newick tree into an
// alignment handle (codes, gap frequencies, rooted pre-order tree), .ct files and
// dot-bracket strings into pairings.
//
//   FASTA + Newick -> getalignmentfromfastaandnewick (src/AlignmentIO.jl:18-48)
//                     and the Dataset constructor (src/Dataset.jl:126-199):
//                     state masks (setnucleotides, src/Dataset.jl:13-96), gap
//                     frequency and observed frequencies (nucmapping lookup,
//                     case-sensitive, src/Dataset.jl:135-152), the tree parsed
//                     (treefromnewickhelper / gettreefromnewick, src/Tree.jl:114-171),
//                     rooted at its first child (roottree, src/Tree.jl:45-55) and
//                     numbered in pre-order (annotatetree, src/Tree.jl:190-199).
//   .ct            -> readctfile (src/RNATools.jl:165-193)
//   dot-bracket    -> parsedbn (src/RNATools.jl:109-140)
//
// Errors are reported through messi_io_error() (per thread), since these calls
// need no device context.
#include <cctype>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/messi.h"

struct messi_alignment {
  std::vector<std::string> names, seqs;
  int nseq = 0, ncol = 0;
  std::vector<uint8_t> codes;  // nseq x ncol state masks, row-major
  std::vector<double> gapfreq;
  double obsfreqs[4] = {0, 0, 0, 0};
  int nnodes = 0;
  std::vector<int> children;  // 2 per node, -1 for leaves
  std::vector<int> seqindex;  // alignment row of a leaf, -1 for internal nodes
  std::vector<double> bl;     // branch length above each node (root: 0)
  std::vector<std::string> nodenames;
};

namespace {

thread_local std::string io_err;

template <class F> int io_guard(F &&f) {
  try {
    f();
    io_err.clear();
    return 0;
  } catch (const std::exception &e) {
    io_err = e.what();
    return 1;
  }
}

std::string read_file(const char *path) {
  std::ifstream in(path, std::ios::binary);
  if (!in) throw std::runtime_error(std::string("cannot open ") + path);
  std::ostringstream ss;
  ss << in.rdbuf();
  return ss.str();
}

std::string strip(const std::string &s) {
  size_t b = 0, e = s.size();
  while (b < e && std::isspace((unsigned char)s[b])) ++b;
  while (e > b && std::isspace((unsigned char)s[e - 1])) --e;
  return s.substr(b, e - b);
}

// 4-bit state mask of an alignment character (A=1, C=2, G=4, T/U=8; IUPAC codes
// set several bits; anything else, gaps included, sets all four).
uint8_t state_mask(char ch) {
  switch (std::toupper((unsigned char)ch)) {
    case 'A': return 0x1;
    case 'C': return 0x2;
    case 'G': return 0x4;
    case 'T': case 'U': return 0x8;
    case 'B': return 0xE;
    case 'D': return 0xD;
    case 'H': return 0xB;
    case 'K': return 0xC;
    case 'M': return 0x3;
    case 'R': return 0x5;
    case 'S': return 0x6;
    case 'V': return 0x7;
    case 'W': return 0x9;
    case 'Y': return 0xA;
    default: return 0xF;
  }
}

struct Node {
  std::string name;
  double bl = 0.0;
  std::vector<std::unique_ptr<Node>> kids;
};

// treefromnewickhelper: the tag after the last ')' is "name:length"; the text
// between the first character and that ')' splits at top-level commas.  Child
// strings are stripped first, so whitespace after a comma does not shift the
// first character off an opening parenthesis.
std::unique_ptr<Node> parse_newick(const std::string &s) {
  auto node = std::make_unique<Node>();
  const size_t close = s.rfind(')');
  std::string tag = s;
  if (close != std::string::npos) {
    tag = s.substr(close + 1);
    const std::string inner = close >= 1 ? s.substr(1, close - 1) : std::string();
    std::string cur;
    int depth = 0;
    for (char c : inner) {
      if (c == '(') ++depth;
      else if (c == ')') --depth;
      if (c == ',' && depth == 0) {
        node->kids.push_back(parse_newick(strip(cur)));
        cur.clear();
      } else {
        cur.push_back(c);
      }
    }
    if (!cur.empty()) node->kids.push_back(parse_newick(strip(cur)));
  }
  const size_t colon = tag.find(':');
  node->name = strip(tag.substr(0, colon));
  if (colon != std::string::npos) {
    std::string len = tag.substr(colon + 1);
    const size_t c2 = len.find(':');
    if (c2 != std::string::npos) len = len.substr(0, c2);
    len = strip(len);
    if (!len.empty()) {
      size_t used = 0;
      try {
        node->bl = std::stod(len, &used);
      } catch (const std::exception &) {
        used = std::string::npos;
      }
      if (used != len.size()) throw std::runtime_error("bad branch length '" + len + "'");
    }
  }
  return node;
}

void preorder(const Node *n, messi_alignment &a, const std::unordered_map<std::string, int> &rows) {
  const int me = a.nnodes++;
  a.children.push_back(-1);
  a.children.push_back(-1);
  a.bl.push_back(n->bl);
  a.nodenames.push_back(n->name);
  if (n->kids.empty()) {
    auto it = rows.find(n->name);
    if (it == rows.end()) throw std::runtime_error("tree leaf '" + n->name + "' has no sequence");
    a.seqindex.push_back(it->second);
    return;
  }
  a.seqindex.push_back(-1);
  if (n->kids.size() != 2)
    throw std::runtime_error("the rooted tree must be binary (the Newick root must be trifurcating)");
  for (int c = 0; c < 2; ++c) {
    a.children[2 * me + c] = a.nnodes;
    preorder(n->kids[c].get(), a, rows);
  }
}

}  // namespace

extern "C" {

const char *messi_io_error(void) { return io_err.c_str(); }

int messi_alignment_read(const char *fasta_path, const char *newick_path, messi_alignment **out) {
  if (!out) return 2;
  *out = nullptr;
  return io_guard([&] {
    if (!fasta_path || !newick_path) throw std::invalid_argument("null path");
    auto a = std::make_unique<messi_alignment>();
    // FASTA: '>' starts a record (its name is the rest of the line, stripped);
    // sequence lines are concatenated without whitespace.
    std::istringstream fa(read_file(fasta_path));
    std::string line;
    while (std::getline(fa, line)) {
      if (!line.empty() && line.back() == '\r') line.pop_back();
      if (!line.empty() && line[0] == '>') {
        a->names.push_back(strip(line.substr(1)));
        a->seqs.emplace_back();
      } else if (!a->seqs.empty()) {
        for (char c : line)
          if (!std::isspace((unsigned char)c)) a->seqs.back().push_back(c);
      }
    }
    if (a->seqs.empty()) throw std::runtime_error("no sequences in FASTA file");
    a->nseq = (int)a->seqs.size();
    a->ncol = (int)a->seqs[0].size();
    for (const auto &s : a->seqs)
      if ((int)s.size() != a->ncol) throw std::runtime_error("aligned sequences must have equal length");
    a->codes.resize((size_t)a->nseq * a->ncol);
    a->gapfreq.assign(a->ncol, 0.0);
    for (int s = 0; s < a->nseq; ++s)
      for (int j = 0; j < a->ncol; ++j) {
        const char ch = a->seqs[s][j];
        a->codes[(size_t)s * a->ncol + j] = state_mask(ch);
        const int k = ch == 'A' ? 0 : ch == 'C' ? 1 : ch == 'G' ? 2 : (ch == 'T' || ch == 'U') ? 3 : -1;
        if (k >= 0) a->obsfreqs[k] += 1.0;
        else a->gapfreq[j] += 1.0;
      }
    double tot = a->obsfreqs[0] + a->obsfreqs[1] + a->obsfreqs[2] + a->obsfreqs[3];
    for (double &f : a->obsfreqs) f = tot > 0 ? f / tot : 0.0;
    for (double &g : a->gapfreq) g /= a->nseq;
    std::unordered_map<std::string, int> rows;
    for (int s = 0; s < a->nseq; ++s) rows[a->names[s]] = s;  // a repeated name keeps its last row
    // Newick: the first line of the file, trailing ';' removed.
    std::istringstream nw(read_file(newick_path));
    std::string nl;
    std::getline(nw, nl);
    nl = strip(nl);
    while (!nl.empty() && nl.back() == ';') nl.pop_back();
    if (nl.empty()) throw std::runtime_error("empty Newick string");
    std::unique_ptr<Node> root = parse_newick(nl);
    // roottree: splice the first child under a new root, halving its branch
    if (root->kids.empty()) throw std::runtime_error("the tree has a single node");
    auto nr = std::make_unique<Node>();
    std::unique_ptr<Node> child = std::move(root->kids.front());
    root->kids.erase(root->kids.begin());
    const double br = child->bl / 2.0;
    child->bl = br;
    root->bl = br;
    nr->kids.push_back(std::move(child));
    nr->kids.push_back(std::move(root));
    preorder(nr.get(), *a, rows);
    *out = a.release();
  });
}

void messi_alignment_destroy(messi_alignment *a) { delete a; }

int messi_alignment_info(const messi_alignment *a, int *nseq, int *ncol, int *nnodes) {
  if (!a) return 2;
  if (nseq) *nseq = a->nseq;
  if (ncol) *ncol = a->ncol;
  if (nnodes) *nnodes = a->nnodes;
  return 0;
}

int messi_alignment_data(const messi_alignment *a, double *data, uint8_t *codes, double *gapfrequency,
                         double *obsfreqs) {
  if (!a) return 2;
  const size_t n = (size_t)a->nseq * a->ncol;
  if (data)  // Julia layout (nseq, ncol, 4), column-major
    for (int s = 0; s < a->nseq; ++s)
      for (int j = 0; j < a->ncol; ++j)
        for (int k = 0; k < 4; ++k)
          data[s + (size_t)a->nseq * (j + (size_t)a->ncol * k)] = (a->codes[(size_t)s * a->ncol + j] >> k) & 1;
  if (codes) std::memcpy(codes, a->codes.data(), n);
  if (gapfrequency) std::memcpy(gapfrequency, a->gapfreq.data(), sizeof(double) * a->ncol);
  if (obsfreqs) std::memcpy(obsfreqs, a->obsfreqs, sizeof(double) * 4);
  return 0;
}

int messi_alignment_tree(const messi_alignment *a, int *children, int *seqindex, double *branchlengths) {
  if (!a) return 2;
  if (children) std::memcpy(children, a->children.data(), sizeof(int) * 2 * a->nnodes);
  if (seqindex) std::memcpy(seqindex, a->seqindex.data(), sizeof(int) * a->nnodes);
  if (branchlengths) std::memcpy(branchlengths, a->bl.data(), sizeof(double) * a->nnodes);
  return 0;
}

int messi_alignment_name(const messi_alignment *a, int row, char *buf, int cap) {
  if (!a || row < 0 || row >= a->nseq || !buf || cap < 1) return 2;
  std::snprintf(buf, (size_t)cap, "%s", a->names[row].c_str());
  return (int)a->names[row].size() >= cap ? 3 : 0;
}

int messi_dataset_from_alignment(messi_ctx *ctx, const messi_alignment *a, int maxbasepairdistance,
                                 const uint8_t *mask, messi_dataset **out) {
  if (!ctx || !a || !out) return 2;
  std::vector<double> data((size_t)a->nseq * a->ncol * 4);
  messi_alignment_data(a, data.data(), nullptr, nullptr, nullptr);
  return messi_dataset_create(ctx, a->nseq, a->ncol, data.data(), a->nnodes, a->children.data(),
                              a->seqindex.data(), maxbasepairdistance, mask, out);
}

int messi_read_ct(const char *path, int cap, int *L, int *paired, char *sequence) {
  if (!L) return 2;
  return io_guard([&] {
    if (!path) throw std::invalid_argument("null path");
    std::istringstream in(read_file(path));
    std::string line;
    if (!std::getline(in, line)) throw std::runtime_error("empty .ct file");
    std::istringstream h(line);
    int len = 0;
    if (!(h >> len) || len < 0) throw std::runtime_error("bad .ct header");
    *L = len;
    if (!paired) return;  // length query
    if (len > cap) throw std::invalid_argument("buffer too small for the .ct length");
    std::vector<int> p(len, 0);
    std::string seq;
    while (std::getline(in, line)) {
      std::istringstream f(line);
      std::vector<std::string> tok;
      std::string t;
      while (f >> t) tok.push_back(t);
      if (tok.size() < 5) continue;
      const int x = std::stoi(tok[0]), y = std::stoi(tok[4]);
      seq += tok[1];
      if (y > x) {
        if (x < 1 || y > len) throw std::runtime_error("pair outside the .ct length");
        if (p[x - 1] == 0 && p[y - 1] == 0) {  // a position already paired keeps its first partner
          p[x - 1] = y;
          p[y - 1] = x;
        }
      }
    }
    std::memcpy(paired, p.data(), sizeof(int) * len);
    if (sequence) std::snprintf(sequence, (size_t)cap + 1, "%s", seq.c_str());
  });
}

int messi_parse_dbn(const char *dbn, int cap, int *L, int *paired) {
  if (!L) return 2;
  return io_guard([&] {
    if (!dbn) throw std::invalid_argument("null string");
    const std::string s = strip(dbn);
    *L = (int)s.size();
    if (!paired) return;
    if ((int)s.size() > cap) throw std::invalid_argument("buffer too small for the structure");
    static const char left[] = "(<ABCDE", right[] = ")>abcde";
    std::vector<int> p(s.size(), 0), st, type;
    for (size_t i = 0; i < s.size(); ++i) {
      const char *l = std::strchr(left, s[i]), *r = std::strchr(right, s[i]);
      if (s[i] && l) {
        st.push_back((int)i);
        type.push_back((int)(l - left));
      } else if (s[i] && r) {
        // only the innermost open bracket of the same kind closes; others are ignored
        if (!type.empty() && type.back() == (int)(r - right)) {
          const int j = st.back();
          st.pop_back();
          type.pop_back();
          p[i] = j + 1;
          p[j] = (int)i + 1;
        }
      }
    }
    std::memcpy(paired, p.data(), sizeof(int) * s.size());
  });
}

}  // extern "C"
