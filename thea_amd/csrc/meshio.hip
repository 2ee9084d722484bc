// Mesh and point-sample files (host code): Thea's OFF and OBJ readers (Graphics/MeshCodecOff.hpp:
// 146-430, Graphics/MeshCodecObj.hpp:210-530) with the options Tools/MeshSample uses, Thea's
// triangulation of the faces read (Algorithms/MeshTriangles.hpp:173-228: triangles as they are,
// quads by Polygon3::triangulateQuad, larger polygons by Polygon3::triangulate), and MeshSample's
// point-file writer and reader (Tools/MeshSample/MeshSample.cpp:314-348, :560-601).
//
// The result is one flat triangle list over the file's vertices, in the order Thea's MeshTriangles
// visits a MeshGroup read from the file (meshes in file order, faces in file order), with the face
// index Thea stores for each face (store_face_indices: the face's position in the file, counting
// faces that were later rejected for having fewer than three vertices).
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <limits>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "common.hpp"

namespace thg {
namespace {

struct V3f {
  float x, y, z;
};
inline V3f sub(V3f a, V3f b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3f cross(V3f a, V3f b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
// 3-vector reductions in Eigen's fixed-size order: x0 + (x1 + x2).
inline float dot(V3f a, V3f b) { return a.x * b.x + (a.y * b.y + a.z * b.z); }
inline float sqnorm(V3f a) { return dot(a, a); }
inline V3f normalized(V3f a) {
  const float n = std::sqrt(sqnorm(a));
  return {a.x / n, a.y / n, a.z / n};
}

struct FileMesh {
  std::vector<float> verts;                  // 3 per vertex
  std::vector<std::vector<int64_t>> faces;   // vertex indices of each accepted face
  std::vector<int64_t> face_index;           // Thea's stored face index of each accepted face
};

bool is_space(char c) { return std::isspace((unsigned char)c) != 0; }

std::string trim(const std::string &s) {
  size_t a = 0, b = s.size();
  while (a < b && is_space(s[a])) ++a;
  while (b > a && is_space(s[b - 1])) --b;
  return s.substr(a, b - a);
}

bool begins_with(const std::string &s, const char *p) { return s.compare(0, std::strlen(p), p) == 0; }

// A byte reader with BinaryInputStream::readLine's newline rule: a line ends at '\r' or '\n', and a
// following newline character different from the first is consumed with it.
struct Reader {
  std::string buf;
  size_t pos = 0;
  bool more() const { return pos < buf.size(); }
  std::string line() {
    size_t e = pos;
    while (e < buf.size() && buf[e] != '\n' && buf[e] != '\r') ++e;
    std::string s = buf.substr(pos, e - pos);
    pos = e;
    if (pos < buf.size()) {
      const char first = buf[pos++];
      if (pos < buf.size() && (buf[pos] == '\n' || buf[pos] == '\r') && buf[pos] != first) ++pos;
    }
    return s;
  }
  std::string content_line(const char *what) {  // the next line that is not empty or a comment
    std::string l = trim(line());
    while (l.empty() || l[0] == '#') {
      if (!more()) fail(THG_E_INVALID, std::string("OFF: unexpected end of input reading ") + what);
      l = trim(line());
    }
    return l;
  }
  int32_t be_i32() {
    if (pos + 4 > buf.size()) fail(THG_E_INVALID, "OFF: unexpected end of binary input");
    const uint32_t v = ((uint32_t)(uint8_t)buf[pos] << 24) | ((uint32_t)(uint8_t)buf[pos + 1] << 16) |
                       ((uint32_t)(uint8_t)buf[pos + 2] << 8) | (uint32_t)(uint8_t)buf[pos + 3];
    pos += 4;
    int32_t r;
    std::memcpy(&r, &v, 4);
    return r;
  }
  float be_f32() {
    const int32_t i = be_i32();
    float f;
    std::memcpy(&f, &i, 4);
    return f;
  }
};

// MeshCodecOff.hpp:190-286 (ASCII). Faces with out-of-range or repeated vertices are skipped (the
// reader's default, non-strict mode); their index is still consumed.
void read_off_ascii(Reader &in, FileMesh &m) {
  const std::string counts_line = in.content_line("the counts");
  std::istringstream counts(counts_line);
  int64_t nv, nf, ne;
  if (!(counts >> nv >> nf >> ne)) fail(THG_E_INVALID, "OFF: expected vertex, face and edge counts, got '" + counts_line + "'");
  if (nv <= 0) return;  // skip_empty_meshes
  for (int64_t v = 0; v < nv; ++v) {
    const std::string l = in.content_line("a vertex");
    std::istringstream vs(l);
    double x, y, z;
    if (!(vs >> x >> y >> z)) fail(THG_E_INVALID, "OFF: expected three vertex coordinates, got '" + l + "'");
    m.verts.push_back((float)x);
    m.verts.push_back((float)y);
    m.verts.push_back((float)z);
  }
  for (int64_t f = 0; f < nf; ++f) {
    const std::string l = in.content_line("a face");
    std::istringstream fs(l);
    int64_t k;
    if (!(fs >> k)) fail(THG_E_INVALID, "OFF: expected a face's vertex count, got '" + l + "'");
    if (k <= 0) continue;
    std::vector<int64_t> face;
    bool skip = false;
    for (int64_t j = 0; j < k && !skip; ++j) {
      int64_t idx;
      if (!(fs >> idx) || idx < 0 || idx >= nv) {
        skip = true;
        break;
      }
      for (int64_t w : face)
        if (w == idx) skip = true;
      face.push_back(idx);
    }
    if (skip || face.size() < 3) continue;  // GeneralMesh::initFace rejects < 3 vertices
    m.faces.push_back(face);
    m.face_index.push_back(f);
  }
}

// MeshCodecOff.hpp:290-430 (binary, big-endian). As in Thea, an out-of-range index ends the face
// early, so the next word is taken as the face's colour count.
void read_off_binary(Reader &in, FileMesh &m) {
  const int64_t nv = in.be_i32(), nf = in.be_i32();
  (void)in.be_i32();  // edges
  if (nv <= 0) return;
  for (int64_t v = 0; v < 3 * nv; ++v) m.verts.push_back(in.be_f32());
  for (int64_t f = 0; f < nf; ++f) {
    const int32_t k = in.be_i32();
    if (k <= 0) continue;
    std::vector<int64_t> face;
    bool skip = false;
    for (int32_t j = 0; j < k; ++j) {
      if (skip) {  // after a repeated vertex the rest of the face is skipped
        (void)in.be_i32();
        continue;
      }
      const int64_t idx = in.be_i32();
      if (idx < 0 || idx >= nv) {  // Thea stops reading the face here (MeshCodecOff.hpp:389-400)
        skip = true;
        break;
      }
      for (int64_t w : face)
        if (w == idx) skip = true;
      face.push_back(idx);
    }
    const int32_t ncolor = in.be_i32();
    if (ncolor > 0) {
      if ((uint64_t)ncolor * 4 > in.buf.size() - in.pos) fail(THG_E_INVALID, "OFF: unexpected end of binary input");
      in.pos += (size_t)ncolor * 4;
    }
    if (skip || face.size() < 3) continue;
    m.faces.push_back(face);
    m.face_index.push_back(f);
  }
}

void read_off(Reader &in, FileMesh &m) {
  const std::string header = trim(in.line());
  if (header != "OFF" && !begins_with(header, "OFF "))
    fail(THG_E_INVALID, "OFF: invalid OFF stream (does not start with 'OFF')");
  if (header == "OFF BINARY" || begins_with(header, "OFF BINARY ")) read_off_binary(in, m);
  else read_off_ascii(in, m);
}

// MeshCodecObj.hpp:210-455 with texture coordinates and normals not read (MeshSample.cpp:493):
// vertex references keep their first integer, negative references count back from the vertices
// read so far, malformed faces are skipped without an index, every other face takes the next
// index (faces of fewer than three vertices are then rejected by the mesh).
void read_obj(Reader &in, FileMesh &m) {
  int64_t num_faces = 0;
  while (in.more()) {
    const std::string l = trim(in.line());
    if (l.empty()) continue;
    const char c = l[0];
    if (c == 'v' && l.size() >= 2 && (l[1] == ' ' || l[1] == '\t')) {
      std::istringstream vs(l.substr(1));
      double x, y, z;
      if (!(vs >> x >> y >> z)) fail(THG_E_INVALID, "OBJ: expected three coordinates after 'v', got '" + l + "'");
      m.verts.push_back((float)x);
      m.verts.push_back((float)y);
      m.verts.push_back((float)z);
    } else if ((c == 'f' || c == 'p') && l.size() >= 2 && (l[1] == ' ' || l[1] == '\t')) {
      const int64_t nverts = (int64_t)(m.verts.size() / 3);
      std::vector<int64_t> face;
      bool bad = false;
      size_t begin = l.find_first_not_of("fp \t"), end = 0;
      while (end != std::string::npos) {
        end = l.find_first_of(" \t", begin);
        std::istringstream fs(l.substr(begin, end == std::string::npos ? std::string::npos : end - begin));
        int64_t v;
        if (!(fs >> v) || std::llabs(v) < 1 || std::llabs(v) > nverts) {
          bad = true;
          break;
        }
        if (v < 0) v = nverts + 1 + v;
        face.push_back(v - 1);
        if (end != std::string::npos) begin = l.find_first_not_of(" \t", end);
      }
      if (bad) continue;
      const int64_t index = num_faces++;
      if (face.size() < 3) continue;
      m.faces.push_back(face);
      m.face_index.push_back(index);
    }
    // 'g'/'o' start a new submesh in Thea; the flat triangle list is the same. Other lines are ignored.
  }
}

// Polygon3::triangulateQuad (Polygon3.hpp:245-289): split along p0-p2 unless the two triangles'
// normals disagree, then along p1-p3; drop triangles whose normal is shorter than float epsilon.
int triangulate_quad(const V3f p[4], int tri[2][3]) {
  V3f n0 = cross(sub(p[1], p[0]), sub(p[2], p[0]));
  V3f n1 = cross(sub(p[2], p[0]), sub(p[3], p[0]));
  int a[2][3] = {{0, 1, 2}, {2, 3, 0}};
  if (dot(n0, n1) < 0) {
    n0 = cross(sub(p[2], p[1]), sub(p[3], p[1]));
    n1 = cross(sub(p[3], p[1]), sub(p[0], p[1]));
    const int b[2][3] = {{1, 2, 3}, {3, 0, 1}};
    std::memcpy(a, b, sizeof(a));
  }
  const float eps = std::numeric_limits<float>::epsilon();
  const float e2 = eps * eps;
  int count = 2;
  if (sqnorm(n0) < e2) {
    if (sqnorm(n1) < e2) return 0;
    std::memcpy(a[0], a[1], sizeof(a[0]));
    count = 1;
  } else if (sqnorm(n1) < e2) {
    count = 1;
  }
  std::memcpy(tri, a, sizeof(a));
  return count;
}

// Polygon3::computeNormal (Polygon3.hpp:180-208): the first corner (p2 - p1) x (p0 - p1) of squared
// length >= 1e-10, normalised; zero if none.
V3f polygon_normal(const std::vector<V3f> &p) {
  const size_t n = p.size();
  V3f p0 = p[0], p1 = p[1 % n];
  size_t i2 = 2 % n;
  for (size_t k = 0; k < n; ++k) {
    const V3f p2 = p[i2];
    const V3f nn = cross(sub(p2, p1), sub(p0, p1));
    if (sqnorm(nn) >= 1e-10f) return normalized(nn);
    p0 = p1;
    p1 = p2;
    i2 = (i2 + 1) % n;
  }
  return {0, 0, 0};
}

struct V2f {
  float x, y;
};

// Polygon3.cpp (insideTriangle2): P inside or on triangle ABC (counter-clockwise).
bool inside_triangle2(V2f A, V2f B, V2f C, V2f P) {
  const float ax = C.x - B.x, ay = C.y - B.y, bx = A.x - C.x, by = A.y - C.y, cx = B.x - A.x, cy = B.y - A.y;
  const float apx = P.x - A.x, apy = P.y - A.y, bpx = P.x - B.x, bpy = P.y - B.y, cpx = P.x - C.x, cpy = P.y - C.y;
  const float a_bp = ax * bpy - ay * bpx, c_ap = cx * apy - cy * apx, b_cp = bx * cpy - by * cpx;
  return a_bp >= 0 && b_cp >= 0 && c_ap >= 0;
}

// Polygon3::triangulate (Polygon3.cpp:178-275): ear clipping (J. Ratcliff's method) in the plane of
// the polygon (orthonormalBasis of its normal, MatrixUtil.hpp:222-246).
void triangulate_polygon(const std::vector<V3f> &p, std::vector<int> &out) {
  const size_t n = p.size();
  const V3f w = polygon_normal(p);
  const V3f wn = normalized(w);
  const float ax_ = std::fabs(w.x), ay_ = std::fabs(w.y), az_ = std::fabs(w.z);
  const int max_axis = (ay_ > ax_) ? ((az_ > ay_) ? 2 : 1) : ((az_ > ax_) ? 2 : 0);
  const V3f u = max_axis == 0 ? normalized({w.y, -w.x, 0.0f}) : normalized({0.0f, w.z, -w.y});
  const V3f v = cross(wn, u);
  std::vector<V2f> q(n);
  for (size_t i = 0; i < n; ++i) {
    const V3f d = sub(p[i], p[0]);
    q[i] = {dot(d, u), dot(d, v)};
  }
  float area = 0;
  for (size_t a = n - 1, b = 0; b < n; a = b++) area += q[a].x * q[b].y - q[b].x * q[a].y;
  area = 0.5f * area;
  std::vector<size_t> idx(n);
  const bool flipped = !(area > 0);
  for (size_t i = 0; i < n; ++i) idx[i] = flipped ? (n - 1) - i : i;
  const float eps = std::numeric_limits<float>::epsilon();
  auto snip = [&](size_t a, size_t b, size_t c, size_t m) {
    const V2f A = q[idx[a]], B = q[idx[b]], C = q[idx[c]];
    if (eps > ((B.x - A.x) * (C.y - A.y)) - ((B.y - A.y) * (C.x - A.x))) return false;
    for (size_t k = 0; k < m; ++k) {
      if (k == a || k == b || k == c) continue;
      if (inside_triangle2(A, B, C, q[idx[k]])) return false;
    }
    return true;
  };
  size_t nv = n, count = 2 * nv;
  for (size_t vv = nv - 1; nv > 2;) {
    if ((count--) <= 0) break;
    size_t a = vv;
    if (nv <= a) a = 0;
    vv = a + 1;
    if (nv <= vv) vv = 0;
    size_t c = vv + 1;
    if (nv <= c) c = 0;
    if (snip(a, vv, c, nv)) {
      const int ia = (int)idx[a], ib = (int)idx[vv], ic = (int)idx[c];
      if (flipped) {
        out.push_back(ic);
        out.push_back(ib);
        out.push_back(ia);
      } else {
        out.push_back(ia);
        out.push_back(ib);
        out.push_back(ic);
      }
      for (size_t s = vv, t = vv + 1; t < nv; ++s, ++t) idx[s] = idx[t];
      --nv;
      count = 2 * nv;
    }
  }
}

// MeshTriangles.hpp:173-228 over every accepted face.
void triangulate(const FileMesh &m, std::vector<int32_t> &tris, std::vector<int64_t> &tri_face) {
  auto pos = [&](int64_t i) { return V3f{m.verts[3 * i], m.verts[3 * i + 1], m.verts[3 * i + 2]}; };
  for (size_t f = 0; f < m.faces.size(); ++f) {
    const std::vector<int64_t> &fv = m.faces[f];
    auto emit = [&](int64_t a, int64_t b, int64_t c) {
      tris.push_back((int32_t)a);
      tris.push_back((int32_t)b);
      tris.push_back((int32_t)c);
      tri_face.push_back(m.face_index[f]);
    };
    if (fv.size() == 3) {
      emit(fv[0], fv[1], fv[2]);
    } else if (fv.size() == 4) {
      const V3f p[4] = {pos(fv[0]), pos(fv[1]), pos(fv[2]), pos(fv[3])};
      int t[2][3];
      const int k = triangulate_quad(p, t);
      for (int j = 0; j < k; ++j) emit(fv[t[j][0]], fv[t[j][1]], fv[t[j][2]]);
    } else {
      std::vector<V3f> p(fv.size());
      for (size_t j = 0; j < fv.size(); ++j) p[j] = pos(fv[j]);
      std::vector<int> t;
      triangulate_polygon(p, t);
      for (size_t j = 0; j + 2 < t.size(); j += 3) emit(fv[t[j]], fv[t[j + 1]], fv[t[j + 2]]);
    }
  }
}

std::string lower_ext(const std::string &path) {
  const size_t dot_pos = path.find_last_of('.');
  std::string e = dot_pos == std::string::npos ? "" : path.substr(dot_pos + 1);
  for (char &ch : e) ch = (char)std::tolower((unsigned char)ch);
  return e;
}

template <typename T> T *copy_out(const std::vector<T> &v) {
  T *p = (T *)std::malloc(std::max<size_t>(1, v.size()) * sizeof(T));
  if (!p) throw std::bad_alloc();
  if (!v.empty()) std::memcpy(p, v.data(), v.size() * sizeof(T));
  return p;
}

template <typename F> int io_guarded(F f) {
  try {
    f();
    return THG_OK;
  } catch (const Status &e) {
    set_last_error(e.msg);
    return e.code;
  } catch (const std::bad_alloc &) {
    set_last_error("host allocation failed");
    return THG_E_OOM;
  } catch (const std::exception &e) {
    set_last_error(e.what());
    return THG_E_INVALID;
  }
}

}  // namespace
}  // namespace thg

using namespace thg;

extern "C" int thg_mesh_file_read(const char *path, thg_mesh_file *out) {
  return io_guarded([&] {
    require(path != nullptr && out != nullptr, "path/out is NULL");
    std::memset(out, 0, sizeof(*out));
    std::ifstream f(path, std::ios::binary);
    if (!f) fail(THG_E_INVALID, std::string("could not open '") + path + "'");
    Reader in;
    in.buf.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
    FileMesh m;
    const std::string ext = lower_ext(path);
    if (ext == "off") read_off(in, m);
    else if (ext == "obj") read_obj(in, m);
    else fail(THG_E_UNSUPPORTED, "unsupported mesh format '." + ext + "' (this build reads .off and .obj)");
    require(m.verts.size() / 3 < (size_t)INT32_MAX, "too many vertices for int32 triangle indices");
    std::vector<int32_t> tris;
    std::vector<int64_t> tri_face;
    triangulate(m, tris, tri_face);
    // Allocate all three arrays before publishing any, so a failed allocation leaks nothing.
    std::unique_ptr<float, decltype(&std::free)> pv(copy_out(m.verts), &std::free);
    std::unique_ptr<int32_t, decltype(&std::free)> pt(copy_out(tris), &std::free);
    std::unique_ptr<int64_t, decltype(&std::free)> pf(copy_out(tri_face), &std::free);
    out->num_vertices = (int64_t)(m.verts.size() / 3);
    out->num_faces = (int64_t)m.faces.size();
    out->num_triangles = (int64_t)tri_face.size();
    out->vertices = pv.release();
    out->triangles = pt.release();
    out->triangle_faces = pf.release();
  });
}

extern "C" void thg_mesh_file_free(thg_mesh_file *m) {
  if (!m) return;
  std::free(m->vertices);
  std::free(m->triangles);
  std::free(m->triangle_faces);
  std::memset(m, 0, sizeof(*m));
}

extern "C" int thg_pts_write(const char *path, int64_t n, const float *pos, const float *nrm, const int64_t *ids) {
  return io_guarded([&] {
    require(path != nullptr && n >= 0 && (n == 0 || pos), "bad path or point array");
    std::ofstream out(path);
    if (!out) fail(THG_E_INVALID, std::string("could not open '") + path + "' for writing");
    // MeshSample.cpp:560-576: values below 1e-35 in magnitude are written as 0.
    auto clean = [](float x) { return std::fabs(x) < 1e-35 ? 0.0f : x; };
    for (int64_t i = 0; i < n; ++i) {
      out << clean(pos[3 * i]) << ' ' << clean(pos[3 * i + 1]) << ' ' << clean(pos[3 * i + 2]);
      if (nrm) out << ' ' << clean(nrm[3 * i]) << ' ' << clean(nrm[3 * i + 1]) << ' ' << clean(nrm[3 * i + 2]);
      if (ids) out << ' ' << ids[i];
      out << '\n';
    }
    out.flush();
    if (!out) fail(THG_E_INVALID, std::string("write to '") + path + "' failed");
  });
}

extern "C" int thg_pts_read(const char *path, int64_t *n, float **pos) {
  return io_guarded([&] {
    require(path != nullptr && n != nullptr && pos != nullptr, "path/n/pos is NULL");
    *n = 0;
    *pos = nullptr;
    std::ifstream in(path);
    if (!in) fail(THG_E_INVALID, std::string("could not open '") + path + "'");
    std::vector<float> p;
    std::string line;
    while (std::getline(in, line)) {  // MeshSample.cpp:328-343
      line = trim(line);
      if (line.empty() || line[0] == '#') continue;
      std::istringstream ls(line);
      float x, y, z;
      if (!(ls >> x >> y >> z)) fail(THG_E_INVALID, "could not read a point from line '" + line + "'");
      p.push_back(x);
      p.push_back(y);
      p.push_back(z);
    }
    *pos = copy_out(p);
    *n = (int64_t)(p.size() / 3);
  });
}

extern "C" void thg_free(void *p) { std::free(p); }
