// OBJ/MTL scene input and BMP image output (host C++), see include/rt_io.h.
//
// The OBJ rules follow what Scene_h::LoadObj gets from the vendored tinyobj
// loader (src/scene.cpp:6-58, inc/tiny_obj_loader.h:1048-1362): number
// syntax and arithmetic of its tryParseDouble (:361-465), index fix-up
// (:312-316), face triples (:523-554), the fan triangulation and face
// grouping of exportFaceGroupToShape (:626-687), and the MTL keys and
// defaults of LoadMtl / InitMaterial (:589-615, :689-1000).
//
// parse_number, fix_index and parse_corner below follow tinyobjloader's
// tryParseDouble, fixIndex and parseTriple step for step (the same
// arithmetic and delimiters, so that vertex and material values come out
// bit-identical), and carry its licence:
//
//   The MIT License (MIT)
//
//   Copyright (c) 2012-2016 Syoyo Fujita and many contributors.
//
//   Permission is hereby granted, free of charge, to any person obtaining a
//   copy of this software and associated documentation files (the
//   "Software"), to deal in the Software without restriction, including
//   without limitation the rights to use, copy, modify, merge, publish,
//   distribute, sublicense, and/or sell copies of the Software, and to
//   permit persons to whom the Software is furnished to do so, subject to
//   the following conditions:
//
//   The above copyright notice and this permission notice shall be included
//   in all copies or substantial portions of the Software.
//
//   THE SOFTWARE IS PROVIDED "AS IS", WITHOUT WARRANTY OF ANY KIND, EXPRESS
//   OR IMPLIED, INCLUDING BUT NOT LIMITED TO THE WARRANTIES OF
//   MERCHANTABILITY, FITNESS FOR A PARTICULAR PURPOSE AND NONINFRINGEMENT.
//   IN NO EVENT SHALL THE AUTHORS OR COPYRIGHT HOLDERS BE LIABLE FOR ANY
//   CLAIM, DAMAGES OR OTHER LIABILITY, WHETHER IN AN ACTION OF CONTRACT,
//   TORT OR OTHERWISE, ARISING FROM, OUT OF OR IN CONNECTION WITH THE
//   SOFTWARE OR THE USE OR OTHER DEALINGS IN THE SOFTWARE.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <fstream>
#include <map>
#include <new>
#include <string>
#include <utility>
#include <vector>

#include "../../include/rt_io.h"

namespace {

inline bool is_space(char c) { return c == ' ' || c == '\t'; }
inline bool is_digit(char c) { return c >= '0' && c <= '9'; }
inline bool is_newline(char c) { return c == '\r' || c == '\n' || c == '\0'; }

// Decimal number in [s, e): sign, integer digits (at least one), optional
// fraction and exponent. The value is accumulated as the loader does:
// mantissa = digits, each fraction digit d at place k adds d * 10^-k, and
// the result is sign * ldexp(mantissa * 5^exp, exp). Returns false (value
// untouched) if there is no leading digit or an exponent has no digits.
bool parse_number(const char* s, const char* e, double* out) {
  if (s >= e) return false;
  const char* c = s;
  double m = 0.0;
  int ex = 0;
  bool neg = false, eneg = false;
  if (*c == '+' || *c == '-') {
    neg = *c == '-';
    ++c;
  } else if (!is_digit(*c)) {
    return false;
  }
  int nd = 0;
  while (c != e && is_digit(*c)) {
    m *= 10;
    m += static_cast<int>(*c - '0');
    ++c;
    ++nd;
  }
  if (nd == 0) return false;
  if (c != e) {
    if (*c == '.') {
      ++c;
      int place = 1;
      while (c != e && is_digit(*c)) {
        m += static_cast<int>(*c - '0') * pow(10.0, -place);
        ++place;
        ++c;
      }
    } else if (!(*c == 'e' || *c == 'E')) {
      c = e;  // anything else ends the number
    }
    if (c != e && (*c == 'e' || *c == 'E')) {
      ++c;
      if (c != e && (*c == '+' || *c == '-')) {
        eneg = *c == '-';
        ++c;
      } else if (c == e || !is_digit(*c)) {
        return false;
      }
      int ne = 0;
      while (c != e && is_digit(*c)) {
        ex = ex * 10 + static_cast<int>(*c - '0');
        ++c;
        ++ne;
      }
      if (eneg) ex = -ex;
      if (ne == 0) return false;
    }
  }
  *out = (neg ? -1 : 1) * ldexp(m * pow(5.0, ex), ex);
  return true;
}

// One whitespace-separated number at *tok (0.0 if it does not parse).
double next_number(const char** tok) {
  *tok += strspn(*tok, " \t");
  const char* end = *tok + strcspn(*tok, " \t\r");
  double v = 0.0;
  parse_number(*tok, end, &v);
  *tok = end;
  return v;
}

void next_numbers(const char** tok, double* v, int k) {
  for (int i = 0; i < k; ++i) v[i] = next_number(tok);
}

std::string first_word(const char* tok) {
  tok += strspn(tok, " \t");
  size_t n = strcspn(tok, " \t\r\n");
  return std::string(tok, n);
}

// 1-based OBJ index -> 0-based; negative = relative to the current count.
int fix_index(int idx, int count) {
  if (idx > 0) return idx - 1;
  if (idx == 0) return 0;
  return count + idx;
}

struct Corner {
  int v = -1, vt = -1, vn = -1;
};

Corner parse_corner(const char** tok, int nv, int nvn, int nvt) {
  Corner c;
  c.v = fix_index(atoi(*tok), nv);
  *tok += strcspn(*tok, "/ \t\r");
  if ((*tok)[0] != '/') return c;
  ++*tok;
  if ((*tok)[0] == '/') {  // v//vn
    ++*tok;
    c.vn = fix_index(atoi(*tok), nvn);
    *tok += strcspn(*tok, "/ \t\r");
    return c;
  }
  c.vt = fix_index(atoi(*tok), nvt);  // v/vt or v/vt/vn
  *tok += strcspn(*tok, "/ \t\r");
  if ((*tok)[0] != '/') return c;
  ++*tok;
  c.vn = fix_index(atoi(*tok), nvn);
  *tok += strcspn(*tok, "/ \t\r");
  return c;
}

// Line reader accepting \n, \r\n and \r endings.
bool read_line(std::istream& in, std::string& line) {
  line.clear();
  if (in.peek() == EOF) return false;
  for (;;) {
    int ch = in.get();
    if (ch == EOF) return true;
    if (ch == '\n') return true;
    if (ch == '\r') {
      if (in.peek() == '\n') in.get();
      return true;
    }
    line.push_back(static_cast<char>(ch));
  }
}

struct MtlEntry {
  std::string name;
  rt_material m;
};

void init_material(MtlEntry& e) {
  e.name.clear();
  memset(&e.m, 0, sizeof(e.m));
  e.m.shininess = 1.0;
}

// MTL file: materials appended in file order; a name already present keeps
// its first id. The pending material is always appended at the end, so a
// missing or empty file still yields one default material.
void load_mtl(std::istream* in, std::vector<MtlEntry>& mats, std::map<std::string, int>& ids) {
  MtlEntry cur;
  init_material(cur);
  std::string line;
  while (in && read_line(*in, line)) {
    size_t last = line.find_last_not_of(" \t");
    line = last == std::string::npos ? std::string() : line.substr(0, last + 1);
    if (line.empty()) continue;
    const char* t = line.c_str();
    t += strspn(t, " \t");
    if (t[0] == '\0' || t[0] == '#') continue;
    if (strncmp(t, "newmtl", 6) == 0 && is_space(t[6])) {
      if (!cur.name.empty()) {
        ids.insert(std::make_pair(cur.name, static_cast<int>(mats.size())));
        mats.push_back(cur);
      }
      init_material(cur);
      cur.name = first_word(t + 7);
      continue;
    }
    double v[3];
    auto key2 = [&](char a, char b) { return t[0] == a && t[1] == b && is_space(t[2]); };
    if (key2('K', 'a')) {
      t += 2;
      next_numbers(&t, v, 3);
      for (int k = 0; k < 3; ++k) cur.m.ka[k] = v[k];
    } else if (key2('K', 'd')) {
      t += 2;
      next_numbers(&t, v, 3);
      for (int k = 0; k < 3; ++k) cur.m.kd[k] = v[k];
    } else if (key2('K', 's')) {
      t += 2;
      next_numbers(&t, v, 3);
      for (int k = 0; k < 3; ++k) cur.m.ks[k] = v[k];
    } else if (key2('K', 't') || key2('T', 'f')) {
      t += 2;
      next_numbers(&t, v, 3);
      for (int k = 0; k < 3; ++k) cur.m.kt[k] = v[k];
    } else if (key2('K', 'e')) {
      t += 2;
      next_numbers(&t, v, 3);
      for (int k = 0; k < 3; ++k) cur.m.ke[k] = v[k];
    } else if (key2('N', 's')) {
      t += 2;
      cur.m.shininess = next_number(&t);
    }
    // other keys (Ni, d, Tr, illum, maps, PBR terms) do not reach the shading
  }
  ids.insert(std::make_pair(cur.name, static_cast<int>(mats.size())));
  mats.push_back(cur);
}

}  // namespace

struct rt_obj {
  std::vector<double> pos64;  // unique vertices, xyz
  std::vector<float> vx, vy, vz, nx, ny, nz;
  std::vector<int32_t> idx, mat_id;
  std::vector<rt_material> mats;
  int32_t nshapes = 0, raw_v = 0, raw_vn = 0;
};

extern "C" {

rt_status rt_obj_load(const char* path, const char* mtl_basepath, rt_obj** out, char* err, size_t errlen) {
  auto fail = [&](rt_status s, const std::string& msg) {
    if (err && errlen) snprintf(err, errlen, "%s", msg.c_str());
    return s;
  };
  if (!out) return fail(RT_ERR_INVALID, "null output");
  *out = nullptr;
  if (!path) return fail(RT_ERR_INVALID, "null path");
  std::ifstream in(path, std::ios::binary);
  if (!in) return fail(RT_ERR_INVALID, std::string("cannot open ") + path);
  const std::string base = mtl_basepath ? mtl_basepath : "";

  std::vector<double> v, vn;
  int nvt = 0;
  std::vector<MtlEntry> mats;
  std::map<std::string, int> mat_ids;
  int material = -1;
  // faces waiting for a flush (corner lists), the current shape's
  // triangles, and the triangles of the shapes already kept
  std::vector<std::vector<Corner>> pending;
  std::vector<Corner> shape_tris;
  std::vector<int> shape_mats;
  std::vector<Corner> tris;
  std::vector<int> tri_mats;
  int nshapes = 0;

  auto flush_pending = [&]() -> bool {
    if (pending.empty()) return false;
    for (const auto& f : pending) {
      for (size_t k = 2; k < f.size(); ++k) {  // fan (0, k-1, k)
        shape_tris.push_back(f[0]);
        shape_tris.push_back(f[k - 1]);
        shape_tris.push_back(f[k]);
        shape_mats.push_back(material);
      }
    }
    return true;
  };
  auto keep_shape = [&]() {
    tris.insert(tris.end(), shape_tris.begin(), shape_tris.end());
    tri_mats.insert(tri_mats.end(), shape_mats.begin(), shape_mats.end());
    ++nshapes;
  };
  auto new_shape = [&]() {
    shape_tris.clear();
    shape_mats.clear();
  };

  std::string line;
  while (read_line(in, line)) {
    if (line.empty()) continue;
    const char* t = line.c_str();
    t += strspn(t, " \t");
    if (t[0] == '\0' || t[0] == '#') continue;
    double xyz[3];
    if (t[0] == 'v' && is_space(t[1])) {
      t += 2;
      next_numbers(&t, xyz, 3);
      v.insert(v.end(), xyz, xyz + 3);
    } else if (t[0] == 'v' && t[1] == 'n' && is_space(t[2])) {
      t += 3;
      next_numbers(&t, xyz, 3);
      vn.insert(vn.end(), xyz, xyz + 3);
    } else if (t[0] == 'v' && t[1] == 't' && is_space(t[2])) {
      t += 3;
      next_numbers(&t, xyz, 2);
      ++nvt;
    } else if (t[0] == 'f' && is_space(t[1])) {
      t += 2;
      t += strspn(t, " \t");
      std::vector<Corner> face;
      while (!is_newline(t[0])) {
        face.push_back(parse_corner(&t, (int)(v.size() / 3), (int)(vn.size() / 3), nvt));
        t += strspn(t, " \t\r");
      }
      pending.push_back(std::move(face));
    } else if (strncmp(t, "usemtl", 6) == 0 && is_space(t[6])) {
      std::string name = first_word(t + 7);
      auto it = mat_ids.find(name);
      int id = it != mat_ids.end() ? it->second : -1;
      if (id != material) {
        flush_pending();  // into the current shape, with the old material
        pending.clear();
        material = id;
      }
    } else if (strncmp(t, "mtllib", 6) == 0 && is_space(t[6])) {
      std::string name = first_word(t + 7);
      std::ifstream mf((base + name).c_str(), std::ios::binary);
      load_mtl(mf ? &mf : nullptr, mats, mat_ids);
    } else if ((t[0] == 'g' || t[0] == 'o') && is_space(t[1])) {
      if (flush_pending()) keep_shape();  // a shape whose pending list is empty is dropped
      new_shape();
      pending.clear();
    }
    // s, t, l, p and unknown lines are ignored
  }
  bool flushed = flush_pending();
  if (flushed || !shape_tris.empty()) keep_shape();

  rt_obj* o = new (std::nothrow) rt_obj();
  if (!o) return fail(RT_ERR_OOM, "host allocation failed");
  o->nshapes = nshapes;
  o->raw_v = (int32_t)(v.size() / 3);
  o->raw_vn = (int32_t)(vn.size() / 3);
  const int nv_raw = o->raw_v, nvn_raw = o->raw_vn;
  std::map<std::pair<int, int>, int> weld;
  o->idx.reserve(tris.size());
  for (const Corner& c : tris) {
    if (c.v < 0 || c.v >= nv_raw || c.vn >= nvn_raw) {
      delete o;
      return fail(RT_ERR_INVALID, "face index out of range");
    }
    auto key = std::make_pair(c.v, c.vn < 0 ? -1 : c.vn);
    auto it = weld.find(key);
    int id;
    if (it == weld.end()) {
      id = (int)o->vx.size();
      weld.emplace(key, id);
      const double* p = &v[3 * (size_t)c.v];
      o->pos64.insert(o->pos64.end(), p, p + 3);
      o->vx.push_back((float)p[0]);
      o->vy.push_back((float)p[1]);
      o->vz.push_back((float)p[2]);
      if (c.vn >= 0) {
        const double* q = &vn[3 * (size_t)c.vn];
        o->nx.push_back((float)q[0]);
        o->ny.push_back((float)q[1]);
        o->nz.push_back((float)q[2]);
      } else {
        o->nx.push_back(0.f);
        o->ny.push_back(0.f);
        o->nz.push_back(0.f);
      }
    } else {
      id = it->second;
    }
    o->idx.push_back(id);
  }
  o->mat_id.assign(tri_mats.begin(), tri_mats.end());
  for (const MtlEntry& e : mats) o->mats.push_back(e.m);
  *out = o;
  return RT_OK;
}

void rt_obj_counts(const rt_obj* o, int32_t* nv, int32_t* n, int32_t* nmat, int32_t* nshapes, int32_t* raw_v,
                   int32_t* raw_vn) {
  if (!o) return;
  if (nv) *nv = (int32_t)o->vx.size();
  if (n) *n = (int32_t)(o->idx.size() / 3);
  if (nmat) *nmat = (int32_t)o->mats.size();
  if (nshapes) *nshapes = o->nshapes;
  if (raw_v) *raw_v = o->raw_v;
  if (raw_vn) *raw_vn = o->raw_vn;
}

void rt_obj_mesh(const rt_obj* o, float* vx, float* vy, float* vz, float* nx, float* ny, float* nz, int32_t* idx,
                 int32_t* mat_id) {
  if (!o) return;
  auto put = [](float* dst, const std::vector<float>& src) {
    if (dst && !src.empty()) memcpy(dst, src.data(), src.size() * sizeof(float));
  };
  put(vx, o->vx);
  put(vy, o->vy);
  put(vz, o->vz);
  put(nx, o->nx);
  put(ny, o->ny);
  put(nz, o->nz);
  if (idx && !o->idx.empty()) memcpy(idx, o->idx.data(), o->idx.size() * sizeof(int32_t));
  if (mat_id && !o->mat_id.empty()) memcpy(mat_id, o->mat_id.data(), o->mat_id.size() * sizeof(int32_t));
}

void rt_obj_positions64(const rt_obj* o, double* xyz) {
  if (o && xyz && !o->pos64.empty()) memcpy(xyz, o->pos64.data(), o->pos64.size() * sizeof(double));
}

void rt_obj_materials(const rt_obj* o, rt_material* mats) {
  if (o && mats && !o->mats.empty()) memcpy(mats, o->mats.data(), o->mats.size() * sizeof(rt_material));
}

void rt_obj_free(rt_obj* o) { delete o; }

rt_status rt_write_bmp(const char* path, int32_t width, int32_t height, const float* rgb) {
  if (!path || !rgb || width <= 0 || height <= 0) return RT_ERR_INVALID;
  const int row = width * 3, pad = (row % 4) ? 4 - (row % 4) : 0, stride = row + pad;
  const uint32_t data_bytes = (uint32_t)stride * (uint32_t)height;
  unsigned char hdr[54];
  memset(hdr, 0, sizeof(hdr));
  auto u16 = [&](int off, uint32_t v) {
    hdr[off] = v & 0xff;
    hdr[off + 1] = (v >> 8) & 0xff;
  };
  auto u32 = [&](int off, uint32_t v) {
    for (int k = 0; k < 4; ++k) hdr[off + k] = (v >> (8 * k)) & 0xff;
  };
  u16(0, 0x4d42);             // "BM"
  // file size field as src/bitmap.cpp:108 computes it: sizeof of the padded
  // 16-byte file-header struct + 40 + pixel bytes (2 more than the file)
  u32(2, 56 + data_bytes);
  u32(10, 54);                // pixel data offset
  u32(14, 40);                // info header size
  u32(18, (uint32_t)width);
  u32(22, (uint32_t)height);  // positive: rows bottom-up
  u16(26, 1);                 // planes
  u16(28, 24);                // bits per pixel
  const uint32_t ppm = (uint32_t)(int)(100 / 2.54 * 72);
  u32(38, ppm);
  u32(42, ppm);
  FILE* f = fopen(path, "wb");
  if (!f) return RT_ERR_INVALID;
  bool ok = fwrite(hdr, 1, sizeof(hdr), f) == sizeof(hdr);
  std::vector<unsigned char> line((size_t)stride, 0);
  for (int j = 0; ok && j < height; ++j) {
    const float* src = rgb + (size_t)j * row;
    for (int i = 0; i < width; ++i) {
      for (int c = 0; c < 3; ++c) {
        double x = src[3 * i + c];
        x = x < 0 ? 0.0 : (x > 1 ? 1.0 : x);
        line[3 * i + (2 - c)] = (unsigned char)(255.0 * x);  // BGR
      }
    }
    ok = fwrite(line.data(), 1, line.size(), f) == line.size();
  }
  ok = (fclose(f) == 0) && ok;
  return ok ? RT_OK : RT_ERR_INVALID;
}

}  // extern "C"
