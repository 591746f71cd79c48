// NASTRAN bulk-data (BDF) mesh reader: the mesh-ingestion step in front of
// the shell path (replaces TACSMeshLoader::scanBDFFile and its accessors,
// include/mesh/TACSMeshLoader.h:145-227, 763-1290, used by
// ElementAssembler::createFromBDF, include/assembler.h:92-115).
//
// Cards read (everything else, and '$' comment lines, is skipped):
//   GRID   small field (8 columns), GRID* large field (16 columns, two
//          lines), and comma-separated free field.
//   SPC    small field: node in columns 16-24, components 24-32, value
//          32-40; SPC* large field: 24-40, 40-56, 56-72.
//   SPC1   SID C G1 G2 ... with continuation lines and "THRU" ranges.
//   CBAR CQUADR CQUAD4 CQUAD8 CQUAD9 CQUAD CHEXA27 CHEXA CTRIA3 CTETRA
//          (a trailing '*' selects 16-column fields): element id, property
//          (component) id, then node ids in columns 8-72; continuation lines
//          start with ' ', '*' or '+' and restart at column 8; a blank field
//          is skipped and a field <= 0 ends the node list.
// The bulk data starts after the last "BEGIN BULK" line (the whole file if
// there is none) and ends at "END BULK" or "ENDDATA".
//
// Numbering: nodes are ordered by their file id, elements by theirs (ids need
// not be contiguous); connectivity and SPC nodes refer to the node order.
// CQUAD4 / CQUADR corner order 1 2 3 4 (around the element) is stored as
// 1 2 4 3 (the tensor-product order of the element), as the reference does.
// Numbers accept the NASTRAN forms 1.5-3 and -2.+4 and a D exponent.
//
// Differences from the reference loader, on purpose: comma-separated GRID
// lines are split at every comma (the reference reads fields from column 8
// on, which is right only when the card name fills columns 0-7); SPC1 is
// read as SPC1 (the reference reads any card starting with "SPC" with the
// SPC layout); columns 72-80 of an element line are its continuation marker
// (the reference reads them as a node field, so a large-field element with a
// marker fails there) and '+' also starts a continuation line; duplicate node
// or element ids are an error; element types
// other than CQUAD4/CQUADR keep their node order (the reference reorders
// CQUAD9, CQUAD and CHEXA as well, and reads past the nodes of CQUAD8 and
// CHEXA27).
#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/gpufem_amd.h"

struct gf_bdf_s {
  std::vector<int> node_ids;      // file ids, ascending
  std::vector<double> xpts;       // 3 per node, in node order
  std::vector<int> elem_ptr{0};   // CSR over elements (ascending file id)
  std::vector<int> conn;          // node indices
  std::vector<int> elem_comp;     // component (property id - 1)
  std::vector<int> bc_nodes, bc_ptr{0}, bc_vars;  // one entry per SPC card
  std::vector<double> bc_vals;
  std::vector<std::string> comp_descript, comp_elem;
};

namespace {

thread_local std::string g_bdf_error;

int fail(const std::string& msg) {
  g_bdf_error = "bdf: " + msg;
  return GF_ERR_INVALID;
}
std::string at_line(size_t k) { return " at line " + std::to_string(k + 1); }

struct ElemType {
  const char* name;
  int min_nodes, max_nodes;
};
// matched in this order by prefix, as the reference does
constexpr ElemType kTypes[] = {{"CBAR", 2, 2},   {"CQUADR", 4, 4},
                               {"CQUAD4", 4, 4}, {"CQUAD8", 8, 8},
                               {"CQUAD9", 9, 9}, {"CQUAD", 9, 9},
                               {"CHEXA27", 27, 27}, {"CHEXA", 8, 8},
                               {"CTRIA3", 3, 3}, {"CTETRA", 4, 10}};

bool starts(const std::string& s, const char* p) {
  return s.compare(0, strlen(p), p) == 0;
}

// columns [a, a + w) of a line (shorter if the line ends first)
std::string field(const std::string& s, size_t a, size_t w) {
  return a < s.size() ? s.substr(a, w) : std::string();
}

bool blank(const std::string& f) {
  for (char c : f)
    if (!isspace((unsigned char)c)) return false;
  return true;
}

int to_int(const std::string& f) { return atoi(f.c_str()); }

// NASTRAN real: 1.0, 1.0E-3, 1.0D-3, and the exponent without a letter
// (1.5-3 = 1.5e-3, -2.+4 = -2e4)
double to_real(std::string f) {
  size_t i = 0;
  while (i < f.size() && isspace((unsigned char)f[i])) i++;
  f = f.substr(i);
  while (!f.empty() && isspace((unsigned char)f.back())) f.pop_back();
  if (f.empty()) return 0.0;
  for (char& c : f)
    if (c == 'd' || c == 'D') c = 'e';
  if (f.find_first_of("eE") == std::string::npos) {
    // a sign after the first character starts the exponent
    for (size_t k = 1; k < f.size(); k++)
      if (f[k] == '+' || f[k] == '-') {
        f.insert(k, 1, 'e');
        break;
      }
  }
  return atof(f.c_str());
}

// comma-separated fields of a free-field card (field 0 is the card name)
std::vector<std::string> split_commas(const std::string& s) {
  std::vector<std::string> out;
  size_t a = 0;
  for (;;) {
    const size_t b = s.find(',', a);
    out.push_back(s.substr(a, b == std::string::npos ? std::string::npos : b - a));
    if (b == std::string::npos) break;
    a = b + 1;
  }
  return out;
}

struct RawElem {
  int id, comp, type;
  std::vector<int> nodes;  // file node ids
};

}  // namespace

extern "C" {

const char* gf_bdf_error(void) { return g_bdf_error.c_str(); }

int gf_bdf_read(const char* path, gf_bdf* out) {
  if (!path || !out) return fail("bad arguments");
  *out = nullptr;
  FILE* fp = fopen(path, "rb");
  if (!fp) return fail(std::string("cannot open ") + path);
  std::string text;
  {
    char buf[1 << 16];
    size_t k;
    while ((k = fread(buf, 1, sizeof buf, fp)) > 0) text.append(buf, k);
    fclose(fp);
  }
  std::vector<std::string> lines;
  for (size_t a = 0; a < text.size();) {
    size_t b = text.find('\n', a);
    if (b == std::string::npos) b = text.size();
    std::string l = text.substr(a, b - a);
    if (!l.empty() && l.back() == '\r') l.pop_back();
    if (l.size() > 80) l.resize(80);
    lines.push_back(std::move(l));
    a = b + 1;
  }
  size_t first = 0;
  for (size_t k = 0; k < lines.size(); k++)
    if (starts(lines[k], "BEGIN BULK")) first = k + 1;

  std::vector<int> nid;
  std::vector<double> nx;
  std::vector<RawElem> elems;
  std::vector<int> bnode, bptr{0}, bvar;
  std::vector<double> bval;
  std::vector<std::string> descript;  // from "$       Shell" comment lines
  auto add_spc = [&](int node, const std::string& comps, double v) {
    bnode.push_back(node);
    for (char c : comps)
      if (c >= '1' && c <= '8') {
        bvar.push_back(c - '1');
        bval.push_back(v);
      }
    bptr.push_back((int)bvar.size());
  };
  for (size_t k = first; k < lines.size(); k++) {
    const std::string& l = lines[k];
    if (starts(l, "$       Shell")) {
      const std::string d = field(l, 41, 32);
      char word[40] = {0};
      if (sscanf(d.c_str(), "%39s", word) != 1) word[0] = 0;
      descript.push_back(word);
      continue;
    }
    if (l.empty() || l[0] == '$') continue;
    if (starts(l, "END BULK") || starts(l, "ENDDATA")) break;
    if (starts(l, "GRID*")) {
      if (k + 1 >= lines.size())
        return fail("GRID*" + at_line(k) + " has no second line");
      const std::string& l2 = lines[++k];
      nid.push_back(to_int(field(l, 8, 16)));
      nx.push_back(to_real(field(l, 40, 16)));
      nx.push_back(to_real(field(l, 56, 16)));
      nx.push_back(to_real(field(l2, 8, 16)));
    } else if (starts(l, "GRID")) {
      const size_t comma = l.find(',');
      if (comma != std::string::npos && comma < 8) {
        // free field: GRID,ID,CP,X1,X2,X3
        const auto f = split_commas(l);
        auto at = [&](size_t q) { return q < f.size() ? f[q] : std::string(); };
        nid.push_back(to_int(at(1)));
        nx.push_back(to_real(at(3)));
        nx.push_back(to_real(at(4)));
        nx.push_back(to_real(at(5)));
      } else if (comma != std::string::npos) {
        // card name in columns 0-7, then comma-separated fields
        const auto f = split_commas(l.substr(8));
        auto at = [&](size_t q) { return q < f.size() ? f[q] : std::string(); };
        nid.push_back(to_int(at(0)));
        nx.push_back(to_real(at(2)));
        nx.push_back(to_real(at(3)));
        nx.push_back(to_real(at(4)));
      } else {
        nid.push_back(to_int(field(l, 8, 8)));
        nx.push_back(to_real(field(l, 24, 8)));
        nx.push_back(to_real(field(l, 32, 8)));
        nx.push_back(to_real(field(l, 40, 8)));
      }
    } else if (starts(l, "SPC1")) {
      // SPC1 SID C G1 ... (small field); "THRU" between two ids is a range
      const std::string comps = field(l, 16, 8);
      std::vector<std::string> f;
      for (size_t a = 24; a < 72 && a < l.size(); a += 8) f.push_back(field(l, a, 8));
      while (k + 1 < lines.size() && !lines[k + 1].empty() &&
             (lines[k + 1][0] == ' ' || lines[k + 1][0] == '+' ||
              lines[k + 1][0] == '*')) {
        const std::string& c = lines[++k];
        for (size_t a = 8; a < 72 && a < c.size(); a += 8) f.push_back(field(c, a, 8));
      }
      std::vector<std::string> t;
      for (auto& s : f)
        if (!blank(s)) t.push_back(s);
      for (size_t q = 0; q < t.size(); q++) {
        std::string u = t[q];
        u.erase(std::remove_if(u.begin(), u.end(),
                               [](char c) { return isspace((unsigned char)c); }),
                u.end());
        for (char& c : u) c = (char)toupper((unsigned char)c);
        if (u == "THRU") {
          if (q == 0 || q + 1 >= t.size())
            return fail("SPC1 THRU without bounds" + at_line(k));
          const int lo = to_int(t[q - 1]), hi = to_int(t[q + 1]);
          for (int g = lo + 1; g <= hi; g++) add_spc(g, comps, 0.0);
          q++;  // hi was added by the range
        } else {
          add_spc(to_int(u), comps, 0.0);
        }
      }
    } else if (starts(l, "SPC*")) {
      add_spc(to_int(field(l, 24, 16)), field(l, 40, 16),
              to_real(field(l, 56, 16)));
    } else if (starts(l, "SPC")) {
      add_spc(to_int(field(l, 16, 8)), field(l, 24, 8), to_real(field(l, 32, 8)));
    } else {
      int type = -1;
      for (int q = 0; q < (int)(sizeof kTypes / sizeof kTypes[0]); q++)
        if (starts(l, kTypes[q].name)) {
          type = q;
          break;
        }
      if (type < 0) continue;  // other cards (properties, materials, loads)
      const ElemType& et = kTypes[type];
      const size_t w = l.size() > strlen(et.name) && l[strlen(et.name)] == '*'
                           ? 16 : 8;
      RawElem e;
      e.type = type;
      e.id = to_int(field(l, 8, w));
      e.comp = to_int(field(l, 8 + w, w));
      if (e.id <= 0 || e.comp <= 0)
        return fail(std::string(et.name) + at_line(k) +
                    " has no positive element and property id");
      size_t at = 8 + 2 * w;
      const std::string* cur = &l;
      bool done = false;
      while ((int)e.nodes.size() < et.max_nodes && !done) {
        // columns 72-80 hold the continuation marker, not a field
        const size_t end = std::min<size_t>(cur->size(), 72);
        for (; (int)e.nodes.size() < et.max_nodes && at < end; at += w) {
          const std::string f = field(*cur, at, w);
          if (blank(f)) continue;
          const int v = to_int(f);
          if (v <= 0) {
            done = true;
            break;
          }
          e.nodes.push_back(v);
        }
        if (done || (int)e.nodes.size() >= et.max_nodes) break;
        // continuation line?
        if (k + 1 < lines.size() && !lines[k + 1].empty() &&
            (lines[k + 1][0] == '*' || lines[k + 1][0] == ' ' ||
             lines[k + 1][0] == '+')) {
          cur = &lines[++k];
          at = 8;
        } else {
          break;
        }
      }
      if ((int)e.nodes.size() < et.min_nodes)
        return fail(std::string(et.name) + at_line(k) + " has too few nodes");
      elems.push_back(std::move(e));
    }
  }

  auto* h = new gf_bdf_s;
  const int nn = (int)nid.size();
  std::vector<int> order(nn);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int b) { return nid[a] < nid[b]; });
  h->node_ids.resize(nn);
  h->xpts.resize(3 * (size_t)nn);
  for (int q = 0; q < nn; q++) {
    h->node_ids[q] = nid[order[q]];
    for (int c = 0; c < 3; c++) h->xpts[3 * q + c] = nx[3 * order[q] + c];
    if (q > 0 && h->node_ids[q] == h->node_ids[q - 1]) {
      const int dup = h->node_ids[q];
      delete h;
      return fail("node id " + std::to_string(dup) + " defined twice");
    }
  }
  auto node_index = [&](int id) {
    auto it = std::lower_bound(h->node_ids.begin(), h->node_ids.end(), id);
    return (it != h->node_ids.end() && *it == id)
               ? (int)(it - h->node_ids.begin()) : -1;
  };
  std::vector<int> eorder(elems.size());
  std::iota(eorder.begin(), eorder.end(), 0);
  std::stable_sort(eorder.begin(), eorder.end(),
                   [&](int a, int b) { return elems[a].id < elems[b].id; });
  int ncomp = 0;  // components: 1 .. largest property id
  for (size_t q = 0; q < eorder.size(); q++) {
    const RawElem& e = elems[eorder[q]];
    if (q > 0 && e.id == elems[eorder[q - 1]].id) {
      delete h;
      return fail("element id " + std::to_string(e.id) + " defined twice");
    }
    std::vector<int> nodes = e.nodes;
    const std::string tn = kTypes[e.type].name;
    if (tn == "CQUAD4" || tn == "CQUADR") std::swap(nodes[2], nodes[3]);
    for (int id : nodes) {
      const int v = node_index(id);
      if (v < 0) {
        delete h;
        return fail("element " + std::to_string(e.id) +
                    " refers to undefined node " + std::to_string(id));
      }
      h->conn.push_back(v);
    }
    h->elem_ptr.push_back((int)h->conn.size());
    h->elem_comp.push_back(e.comp - 1);
    ncomp = std::max(ncomp, e.comp);
  }
  h->comp_descript.assign(ncomp, "");
  h->comp_elem.assign(ncomp, "");
  for (size_t q = 0; q < descript.size() && (int)q < ncomp; q++)
    h->comp_descript[q] = descript[q];
  // a component's element name is that of its first element in file order
  for (const RawElem& e : elems) {
    std::string& s = h->comp_elem[e.comp - 1];
    if (s.empty())
      s = (std::string(kTypes[e.type].name) == "CTETRA" && e.nodes.size() == 10)
              ? "CTETRA10" : kTypes[e.type].name;
  }
  for (size_t q = 0; q < bnode.size(); q++) {
    const int v = node_index(bnode[q]);
    if (v < 0) {
      delete h;
      return fail("SPC on undefined node " + std::to_string(bnode[q]));
    }
    h->bc_nodes.push_back(v);
  }
  h->bc_ptr = bptr;
  h->bc_vars = bvar;
  h->bc_vals = bval;
  *out = h;
  g_bdf_error.clear();
  return GF_OK;
}

int gf_bdf_destroy(gf_bdf h) {
  delete h;
  return GF_OK;
}

int gf_bdf_sizes(gf_bdf h, int* nnodes, int* nelems, int* conn_size,
                 int* nbcs, int* nbc_vars, int* ncomponents) {
  if (!h) return GF_ERR_INVALID;
  if (nnodes) *nnodes = (int)h->node_ids.size();
  if (nelems) *nelems = (int)h->elem_comp.size();
  if (conn_size) *conn_size = (int)h->conn.size();
  if (nbcs) *nbcs = (int)h->bc_nodes.size();
  if (nbc_vars) *nbc_vars = (int)h->bc_vars.size();
  if (ncomponents) *ncomponents = (int)h->comp_elem.size();
  return GF_OK;
}

int gf_bdf_connectivity(gf_bdf h, double* xpts, int* elem_ptr, int* conn,
                        int* elem_comp) {
  if (!h) return GF_ERR_INVALID;
  if (xpts) std::copy(h->xpts.begin(), h->xpts.end(), xpts);
  if (elem_ptr) std::copy(h->elem_ptr.begin(), h->elem_ptr.end(), elem_ptr);
  if (conn) std::copy(h->conn.begin(), h->conn.end(), conn);
  if (elem_comp) std::copy(h->elem_comp.begin(), h->elem_comp.end(), elem_comp);
  return GF_OK;
}

int gf_bdf_bcs(gf_bdf h, int* bc_nodes, int* bc_ptr, int* bc_vars,
               double* bc_vals) {
  if (!h) return GF_ERR_INVALID;
  if (bc_nodes) std::copy(h->bc_nodes.begin(), h->bc_nodes.end(), bc_nodes);
  if (bc_ptr) std::copy(h->bc_ptr.begin(), h->bc_ptr.end(), bc_ptr);
  if (bc_vars) std::copy(h->bc_vars.begin(), h->bc_vars.end(), bc_vars);
  if (bc_vals) std::copy(h->bc_vals.begin(), h->bc_vals.end(), bc_vals);
  return GF_OK;
}

int gf_bdf_assembler_data(gf_bdf h, int vars_per_node, int* elem_conn,
                          int* bcs, double* xpts) {
  if (!h || vars_per_node <= 0) return GF_ERR_INVALID;
  if (elem_conn) std::copy(h->conn.begin(), h->conn.end(), elem_conn);
  if (bcs)
    for (size_t b = 0; b + 1 < h->bc_ptr.size(); b++)
      for (int j = h->bc_ptr[b]; j < h->bc_ptr[b + 1]; j++)
        bcs[j] = vars_per_node * h->bc_nodes[b] + h->bc_vars[j];
  if (xpts) std::copy(h->xpts.begin(), h->xpts.end(), xpts);
  return GF_OK;
}

const char* gf_bdf_component_descript(gf_bdf h, int comp) {
  if (!h || comp < 0 || comp >= (int)h->comp_descript.size()) return nullptr;
  return h->comp_descript[comp].c_str();
}

const char* gf_bdf_element_descript(gf_bdf h, int comp) {
  if (!h || comp < 0 || comp >= (int)h->comp_elem.size()) return nullptr;
  return h->comp_elem[comp].c_str();
}

}  // extern "C"
