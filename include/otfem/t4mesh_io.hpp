// Tetrahedral mesh files for the otfem front end: T4MESH XML read / write and
// the TetGen reader, for any mesh type with the T4Mesh interface of
// <otfem/fem.hpp> (clear(), insert(), insert(i, j, k, m), node(idx),
// tetrahedron iteration, size_nodes()).
//
// Reference interfaces mirrored (paths relative to the OpenTissue tree):
//   t4mesh::xml_read        core/containers/t4mesh/io/t4mesh_xml_read.h:42-200, 250-254
//   t4mesh::xml_write       core/containers/t4mesh/io/t4mesh_xml_write.h:82-115, 134-170
//   t4mesh::tetgen_read     core/containers/t4mesh/io/t4mesh_tetgen_read.h:29-136, 150-154
//
// File format (as the reference reads and writes it):
//   <T4MESH nodes="N">
//     <TETRAHEDRON i="a" j="b" k="c" m="d"/>   (zero-based node indices)
//     <POINT idx="n" coord="[x,y,z]"/>
//   </T4MESH>
// The reference parses it with TinyXML; this header scans the few constructs
// the format uses (elements, quoted attributes, the XML declaration,
// comments) itself, so it needs no XML library.  Optional `tag` attributes
// are accepted and ignored (the otfem mesh types carry no tags).  Like the
// reference, the readers return false (with a message on std::cerr) on a
// missing attribute or an index out of range; where the reference asserts
// (no T4MESH element, a node count that is not positive, a coordinate that is
// not a finite number) these return false as well.  xml_write prints
// coordinates with 17 significant digits (the reference: 15), so a write
// followed by a read reproduces every coordinate exactly.
#ifndef OTFEM_T4MESH_IO_HPP
#define OTFEM_T4MESH_IO_HPP

#include <cmath>
#include <cstddef>
#include <fstream>
#include <iostream>
#include <iterator>
#include <map>
#include <sstream>
#include <string>
#include <utility>
#include <vector>

namespace otfem {
namespace t4mesh {

namespace io_detail {

struct XmlElement {
  std::string name;
  std::vector<std::pair<std::string, std::string>> attrs;
  int depth = 0;  // 0 for a root element
  const std::string *attr(const char *key) const {
    for (const auto &kv : attrs)
      if (kv.first == key) return &kv.second;
    return nullptr;
  }
};

inline std::string xml_unescape(const std::string &s) {
  std::string out;
  for (size_t i = 0; i < s.size(); ++i) {
    if (s[i] != '&') {
      out += s[i];
      continue;
    }
    const size_t e = s.find(';', i);
    if (e == std::string::npos) {
      out += s[i];
      continue;
    }
    const std::string ent = s.substr(i + 1, e - i - 1);
    if (ent == "lt") out += '<';
    else if (ent == "gt") out += '>';
    else if (ent == "amp") out += '&';
    else if (ent == "quot") out += '"';
    else if (ent == "apos") out += '\'';
    else out += s.substr(i, e - i + 1);
    i = e;
  }
  return out;
}

// Start tags of a document in order, with their nesting depth.  Returns false
// on malformed markup.
inline bool xml_scan(const std::string &doc, std::vector<XmlElement> &out) {
  int depth = 0;
  size_t p = 0;
  const size_t n = doc.size();
  auto is_space = [](char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r'; };
  while ((p = doc.find('<', p)) != std::string::npos) {
    if (doc.compare(p, 4, "<!--") == 0) {
      const size_t e = doc.find("-->", p + 4);
      if (e == std::string::npos) return false;
      p = e + 3;
      continue;
    }
    if (doc.compare(p, 2, "<?") == 0 || doc.compare(p, 2, "<!") == 0) {
      const size_t e = doc.find('>', p);
      if (e == std::string::npos) return false;
      p = e + 1;
      continue;
    }
    if (doc.compare(p, 2, "</") == 0) {
      const size_t e = doc.find('>', p);
      if (e == std::string::npos) return false;
      if (--depth < 0) return false;
      p = e + 1;
      continue;
    }
    XmlElement el;
    size_t q = p + 1;
    while (q < n && !is_space(doc[q]) && doc[q] != '/' && doc[q] != '>') el.name += doc[q++];
    if (el.name.empty()) return false;
    bool empty_element = false;
    for (;;) {
      while (q < n && is_space(doc[q])) ++q;
      if (q >= n) return false;
      if (doc[q] == '/') {
        if (q + 1 >= n || doc[q + 1] != '>') return false;
        empty_element = true;
        q += 2;
        break;
      }
      if (doc[q] == '>') {
        ++q;
        break;
      }
      std::string key;
      while (q < n && !is_space(doc[q]) && doc[q] != '=' && doc[q] != '/' && doc[q] != '>')
        key += doc[q++];
      while (q < n && is_space(doc[q])) ++q;
      if (key.empty() || q >= n || doc[q] != '=') return false;
      ++q;
      while (q < n && is_space(doc[q])) ++q;
      if (q >= n || (doc[q] != '"' && doc[q] != '\'')) return false;
      const char quote = doc[q++];
      const size_t e = doc.find(quote, q);
      if (e == std::string::npos) return false;
      el.attrs.emplace_back(key, xml_unescape(doc.substr(q, e - q)));
      q = e + 1;
    }
    el.depth = depth;
    out.push_back(std::move(el));
    if (!empty_element) ++depth;
    p = q;
  }
  return depth == 0;
}

template <typename T>
inline bool parse_value(const std::string &s, T &v) {
  std::istringstream in(s);
  in >> v;
  return !in.fail();
}

// "[x,y,z]" as math_vector3.h:252-263 reads it: one separator character
// before, between and after the three numbers.
inline bool parse_coord(const std::string &s, double v[3]) {
  std::istringstream in(s);
  char dummy;
  in >> dummy >> v[0] >> dummy >> v[1] >> dummy >> v[2];
  return !in.fail();
}

// Whitespace-separated tokens of a TetGen file; '#' starts a comment that
// runs to the end of the line.
inline bool tetgen_tokens(const std::string &filename, std::vector<std::string> &tok) {
  std::ifstream f(filename.c_str());
  if (!f.is_open()) return false;
  std::string line;
  while (std::getline(f, line)) {
    const size_t h = line.find('#');
    if (h != std::string::npos) line.erase(h);
    std::istringstream in(line);
    std::string t;
    while (in >> t) tok.push_back(t);
  }
  return true;
}

}  // namespace io_detail

// t4mesh_xml_read.h:42-200: reads the tetrahedra into `mesh` and the node
// coordinates into `points` (resized to the node count; nodes without a
// POINT element keep a default coordinate).
template <typename point_container, typename t4mesh_type>
bool xml_read(std::string const &filename, t4mesh_type &mesh, point_container &points) {
  mesh.clear();
  std::ifstream f(filename.c_str(), std::ios::binary);
  if (!f.is_open()) {
    std::cerr << "file not found" << std::endl;
    return false;
  }
  const std::string doc((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
  std::vector<io_detail::XmlElement> els;
  if (!io_detail::xml_scan(doc, els)) {
    std::cerr << "t4mesh::xml_read(): malformed XML" << std::endl;
    return false;
  }
  size_t root = els.size();
  for (size_t e = 0; e < els.size(); ++e)
    if (els[e].depth == 0 && els[e].name == "T4MESH") {
      root = e;
      break;
    }
  if (root == els.size()) {
    std::cerr << "t4mesh::xml_read(): no T4MESH element" << std::endl;
    return false;
  }
  long long cnt_nodes = 0;
  if (const std::string *a = els[root].attr("nodes")) io_detail::parse_value(*a, cnt_nodes);
  if (cnt_nodes <= 0) {
    std::cerr << "t4mesh::xml_read(): node count was not positive" << std::endl;
    return false;
  }
  for (long long i = 0; i < cnt_nodes; ++i) mesh.insert();
  // children of the T4MESH element: the elements after it at depth 1, up to
  // the next root element
  size_t end = root + 1;
  while (end < els.size() && els[end].depth > 0) ++end;
  for (size_t e = root + 1; e < end; ++e) {
    const io_detail::XmlElement &el = els[e];
    if (el.depth != 1 || el.name != "TETRAHEDRON") continue;
    long long idx[4] = {-1, -1, -1, -1};
    const char *keys[4] = {"i", "j", "k", "m"};
    for (int c = 0; c < 4; ++c) {
      const std::string *a = el.attr(keys[c]);
      if (!a) {
        std::cerr << "t4mesh::xml_read(): Missing " << keys[c] << " index on tetrahedron"
                  << std::endl;
        return false;
      }
      io_detail::parse_value(*a, idx[c]);
    }
    for (int c = 0; c < 4; ++c)
      if (idx[c] < 0 || idx[c] >= cnt_nodes) {
        std::cerr << "t4mesh::xml_read(): Illegal node index on tetrahedron" << std::endl;
        return false;
      }
    mesh.insert((size_t)idx[0], (size_t)idx[1], (size_t)idx[2], (size_t)idx[3]);
  }
  points.resize((size_t)cnt_nodes);
  for (size_t e = root + 1; e < end; ++e) {
    const io_detail::XmlElement &el = els[e];
    if (el.depth != 1 || el.name != "POINT") continue;
    const std::string *a = el.attr("idx");
    if (!a) {
      std::cerr << "t4mesh::xml_read(): Missing index on point" << std::endl;
      return false;
    }
    long long idx = -1;
    io_detail::parse_value(*a, idx);
    if (idx < 0 || idx >= cnt_nodes) {
      std::cerr << "t4mesh::xml_read(): Illegal index on point" << std::endl;
      return false;
    }
    const std::string *c = el.attr("coord");
    if (!c) {
      std::cerr << "t4mesh::xml_read(): Missing coord on point" << std::endl;
      return false;
    }
    double v[3];
    if (!io_detail::parse_coord(*c, v) || !std::isfinite(v[0]) || !std::isfinite(v[1]) ||
        !std::isfinite(v[2])) {
      std::cerr << "t4mesh::xml_read(): coordinate is not a finite number" << std::endl;
      return false;
    }
    for (int d = 0; d < 3; ++d) points[(size_t)idx](d) = v[d];
  }
  return true;
}

namespace io_detail {
// Coordinates written straight into the nodes (the reference's
// default_point_container, t4mesh_default_point_container.h).
template <typename t4mesh_type>
struct NodeCoords {
  typedef typename t4mesh_type::vector3_type value_type;
  t4mesh_type *mesh;
  void resize(size_t) {}
  size_t size() const { return mesh->size_nodes(); }
  value_type &operator[](size_t i) { return mesh->node(i)->m_coord; }
};
}  // namespace io_detail

// t4mesh_xml_read.h:250-254: coordinates read into node->m_coord.
template <typename t4mesh_type>
bool xml_read(std::string const &filename, t4mesh_type &mesh) {
  io_detail::NodeCoords<t4mesh_type> points{&mesh};
  return xml_read(filename, mesh, points);
}

// t4mesh_xml_write.h:82-115, 134-149: the mesh and `points` (one per node) as
// a T4MESH document.
template <typename point_container, typename t4mesh_type>
bool xml_write(std::string const &filename, t4mesh_type &mesh, point_container &points) {
  if (points.size() != mesh.size_nodes()) {
    std::cerr << "t4mesh::xml_write(): node size mismatch between point container and mesh"
              << std::endl;
    return false;
  }
  std::ofstream f(filename.c_str());
  if (!f.is_open()) return false;
  f << "<?xml version=\"1.0\" ?>\n<T4MESH nodes=\"" << mesh.size_nodes() << "\">\n";
  for (auto t = mesh.tetrahedron_begin(); t != mesh.tetrahedron_end(); ++t)
    f << "    <TETRAHEDRON i=\"" << t->i()->idx() << "\" j=\"" << t->j()->idx() << "\" k=\""
      << t->k()->idx() << "\" m=\"" << t->m()->idx() << "\" />\n";
  f.precision(17);
  for (size_t i = 0; i < mesh.size_nodes(); ++i)
    f << "    <POINT idx=\"" << i << "\" coord=\"[" << points[i](0) << "," << points[i](1) << ","
      << points[i](2) << "]\" />\n";
  f << "</T4MESH>\n";
  return (bool)f;
}

// t4mesh_xml_write.h:163-170: coordinates taken from node->m_coord.
template <typename t4mesh_type>
bool xml_write(std::string const &filename, t4mesh_type &mesh) {
  io_detail::NodeCoords<t4mesh_type> points{&mesh};
  return xml_write(filename, mesh, points);
}

// t4mesh_tetgen_read.h:29-136: `filename`.node and `filename`.ele as TetGen
// writes them (any index base; node attributes, boundary markers and one
// region attribute per tetrahedron are skipped).  Nodes are numbered in file
// order.
template <typename point_container, typename t4mesh_type>
bool tetgen_read(const std::string &filename, t4mesh_type &mesh, point_container &points) {
  mesh.clear();
  std::vector<std::string> tok;
  const std::string node_filename = filename + ".node", element_filename = filename + ".ele";
  if (!io_detail::tetgen_tokens(node_filename, tok)) {
    std::cerr << "Error unable to open file '" << node_filename << "'" << std::endl;
    return false;
  }
  size_t p = 0;
  auto next = [&](auto &v) { return p < tok.size() && io_detail::parse_value(tok[p++], v); };
  long long cnt_nodes = 0, dimension = 0, cnt_attributes = 0, has_boundary = 0;
  if (!next(cnt_nodes) || !next(dimension) || !next(cnt_attributes) || !next(has_boundary) ||
      cnt_nodes < 0 || dimension != 3 || cnt_attributes < 0) {
    std::cerr << "t4mesh::tetgen_read(): bad header in '" << node_filename << "'" << std::endl;
    return false;
  }
  points.resize((size_t)cnt_nodes);
  std::map<long long, size_t> internal;
  for (long long i = 0; i < cnt_nodes; ++i) {
    long long ext;
    double x, y, z, skip;
    if (!next(ext) || !next(x) || !next(y) || !next(z)) {
      std::cerr << "t4mesh::tetgen_read(): '" << node_filename << "' ends early" << std::endl;
      return false;
    }
    for (long long a = 0; a < cnt_attributes + (has_boundary ? 1 : 0); ++a)
      if (!next(skip)) {
        std::cerr << "t4mesh::tetgen_read(): '" << node_filename << "' ends early" << std::endl;
        return false;
      }
    auto node = mesh.insert();
    points[node->idx()](0) = x;
    points[node->idx()](1) = y;
    points[node->idx()](2) = z;
    internal[ext] = node->idx();
  }
  tok.clear();
  p = 0;
  if (!io_detail::tetgen_tokens(element_filename, tok)) {
    std::cerr << "Error unable to open file '" << element_filename << "'" << std::endl;
    return false;
  }
  long long cnt_tetrahedra = 0, cnt_points = 0;
  if (!next(cnt_tetrahedra) || !next(cnt_points) || !next(cnt_attributes) || cnt_tetrahedra < 0) {
    std::cerr << "t4mesh::tetgen_read(): bad header in '" << element_filename << "'" << std::endl;
    return false;
  }
  if (cnt_points != 4) {
    std::cerr << "We only support tetrahedra with 4 points!" << std::endl;
    return false;
  }
  if (cnt_attributes > 1) {
    std::cerr << "We only support tetrahedra with 1 region!" << std::endl;
    return false;
  }
  for (long long i = 0; i < cnt_tetrahedra; ++i) {
    long long ext, n[4];
    double skip;
    if (!next(ext) || !next(n[0]) || !next(n[1]) || !next(n[2]) || !next(n[3]) ||
        (cnt_attributes == 1 && !next(skip))) {
      std::cerr << "t4mesh::tetgen_read(): '" << element_filename << "' ends early" << std::endl;
      return false;
    }
    size_t in[4];
    for (int c = 0; c < 4; ++c) {
      auto it = internal.find(n[c]);
      if (it == internal.end()) {
        std::cerr << "t4mesh::tetgen_read(): tetrahedron " << ext << " names unknown node " << n[c]
                  << std::endl;
        return false;
      }
      in[c] = it->second;
    }
    mesh.insert(in[0], in[1], in[2], in[3]);
  }
  return true;
}

// t4mesh_tetgen_read.h:150-154: coordinates read into node->m_coord.
template <typename t4mesh_type>
bool tetgen_read(const std::string &filename, t4mesh_type &mesh) {
  io_detail::NodeCoords<t4mesh_type> points{&mesh};
  return tetgen_read(filename, mesh, points);
}

}  // namespace t4mesh
}  // namespace otfem

#endif  // OTFEM_T4MESH_IO_HPP
