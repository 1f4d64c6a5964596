// obj_load.cpp — Wavefront OBJ + MTL ingest (obj_load, obj_load.cpp:181-568)
// and PNG decoding for the textures it names; host code of libocr_engine.so.
// The parse follows the reference's rules field for field (see ocr_obj.h),
// so a scene loads into the same triangles, parts and textures.
#include "ocr_obj.h"

#include <zlib.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

struct ocr_obj_mesh {
  std::vector<ocr_triangle> tris;
  std::vector<ocr_obj_part> parts;
};

namespace {

// Lines as the reference reads them: getline while the stream is good, so a
// file ending in '\n' yields a last empty line (obj_load.cpp:223-236).
bool read_lines(const std::string& path, std::vector<std::string>& out) {
  std::ifstream f(path.c_str());
  if (!f.is_open()) return false;
  while (f.good()) {
    std::string s;
    std::getline(f, s);
    out.push_back(s);
  }
  return true;
}

bool file_exists(const std::string& p) {   // obj_load.cpp:18-24
  std::ifstream f(p.c_str());
  return f.good();
}

// std::string::operator[] gives '\0' at size(); the reference reads up to
// three characters of every line.
char at(const std::string& s, size_t i) { return i < s.size() ? s[i] : '\0'; }

// Positions in the reference are ints, so a failed find is -1 and "+ 1"
// restarts from the first character.
int find_int(const std::string& s, char c, int from) {
  const size_t r = s.find(c, (size_t)(from < 0 ? 0 : from));
  return r == std::string::npos ? -1 : (int)r;
}

// decompose_attribute (obj_load.cpp:144-165): field k is read from the
// character after the k-th space; atoi when the line has no '.'.
void attribute(const std::string& str, float* a, int n) {
  const bool has_dot = str.find('.') != std::string::npos;
  int prev = find_int(str, ' ', 0);
  for (int i = 1; i <= n; i++) {
    const int next = find_int(str, ' ', prev + 1);
    const char* p = str.c_str() + (prev + 1);
    a[i - 1] = has_dot ? (float)std::atof(p) : (float)std::atoi(p);
    prev = next;
  }
}

// decompose_face (obj_load.cpp:168-186): three v/vt/vn triples, 1-based.
void face(const std::string& str, int v[3], int vt[3], int vn[3]) {
  int start = 2;
  for (int i = 0; i < 3; i++) {
    const int s1 = find_int(str, '/', start);
    const int s2 = find_int(str, '/', s1 + 1);
    const int s3 = find_int(str, ' ', s2 + 1);
    v[i] = std::atoi(str.c_str() + start) - 1;
    vt[i] = std::atoi(str.c_str() + (s1 + 1)) - 1;
    vn[i] = std::atoi(str.c_str() + (s2 + 1)) - 1;
    start = s3 + 1;
  }
}

std::string after_last_space(const std::string& s) {
  const size_t p = s.find_last_of(' ');
  return s.substr(p == std::string::npos ? 0 : p + 1);
}

// is_samemtl (obj_load.cpp:26-29): "newmtl " and the last word starts with name.
bool same_mtl(const std::string& line, const std::string& name) {
  if (line.compare(0, 7, "newmtl ") != 0) return false;
  const size_t p = line.find_last_of(' ');
  return line.substr(p == std::string::npos ? 0 : p + 1, name.size()) == name;
}

// retrieve_diffuse_new (obj_load.cpp:32-51): the first map_Kd at or after
// the material's newmtl, wherever it is.
std::string diffuse_name(const std::vector<std::string>& mtl, const std::string& name) {
  bool found = false;
  for (const std::string& l : mtl) {
    if (same_mtl(l, name)) found = true;
    if (found && l.compare(0, 7, "map_Kd ") == 0) return after_last_space(l);
  }
  return "";
}

// retrieve_bumpmap (obj_load.cpp:54-79): map_Bump inside the material's block.
std::string bump_name(const std::vector<std::string>& mtl, const std::string& name) {
  bool found = false;
  for (const std::string& l : mtl) {
    if (same_mtl(l, name)) {
      found = true;
      continue;
    }
    if (found && l.compare(0, 9, "map_Bump ") == 0) return after_last_space(l);
    if (found && l.compare(0, 7, "newmtl ") == 0) return "None";
  }
  return "None";
}

// get_Kd (obj_load.cpp:83-119): the first line beginning "Kd" after the
// material's newmtl, fields split on single spaces.
bool kd_of(const std::vector<std::string>& mtl, const std::string& name, float kd[3]) {
  bool found = false;
  for (const std::string& l : mtl) {
    if (same_mtl(l, name)) {
      found = true;
      continue;
    }
    if (!found || l.compare(0, 2, "Kd") != 0) continue;
    std::vector<std::string> f;
    std::stringstream ss(l);
    std::string item;
    while (std::getline(ss, item, ' ')) f.push_back(item);
    if (f.size() <= 3) return false;   // the reference asserts here
    for (int c = 0; c < 3; c++) kd[c] = (float)std::atof(f[1 + c].c_str());
    return true;
  }
  return false;
}

// col2cachename (texture.cpp:9-12): the cache key of a colour texture.
std::string colour_key(float r, float g, float b) {
  return std::string("CACHE") + std::to_string(r) + std::to_string(g) + std::to_string(b) + std::to_string(255.f);
}

void copy_str(char* dst, size_t cap, const std::string& s) {
  std::strncpy(dst, s.c_str(), cap - 1);
  dst[cap - 1] = '\0';
}

uint8_t to_u8(float v) {   // sf::Color(Uint8) from the float Kd * 255
  if (!(v > 0.f)) return 0;
  return v >= 255.f ? 255 : (uint8_t)v;
}

// ---------------------------------------------------------------- PNG
uint32_t be32(const uint8_t* p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }

int paeth(int a, int b, int c) {
  const int p = a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
  return (pa <= pb && pa <= pc) ? a : (pb <= pc ? b : c);
}

}  // namespace

extern "C" {

int ocr_obj_load(const char* path, float scale, ocr_obj_mesh** out) {
  if (!path || !out) return OCR_ERR_INVALID;
  *out = nullptr;
  const std::string filename(path);
  const size_t dotp = filename.find_last_of('.');
  const std::string mtlname = filename.substr(0, dotp) + ".mtl";
  const size_t lslash = filename.find_last_of('/');
  const std::string dir = lslash == std::string::npos ? std::string(".") : filename.substr(0, lslash);
  std::vector<std::string> lines, mtl;
  if (!read_lines(filename, lines)) return OCR_ERR_INVALID;
  read_lines(mtlname, mtl);   // a missing .mtl leaves every part magenta

  struct Idx { int v[3], vt[3], vn[3]; };
  std::vector<float> vl, vtl, vnl;
  std::vector<Idx> fl;
  std::vector<uint32_t> use_pos, group_pos;
  std::vector<std::string> use_name, group_name;
  uint32_t nface = 0;
  for (const std::string& l : lines) {
    const char c0 = at(l, 0), c1 = at(l, 1), c2 = at(l, 2);
    if (c0 == 'f' && c1 == ' ') {
      Idx f;
      face(l, f.v, f.vt, f.vn);
      fl.push_back(f);
      nface++;
    } else if (c0 == 'v' && c1 == ' ') {
      float a[3];
      attribute(l, a, 3);
      vl.insert(vl.end(), a, a + 3);
    } else if (c0 == 'v' && c1 == 't' && c2 == ' ') {
      float a[2];
      attribute(l, a, 2);
      vtl.insert(vtl.end(), a, a + 2);
    } else if (c0 == 'v' && c1 == 'n' && c2 == ' ') {
      float a[3];
      attribute(l, a, 3);
      vnl.insert(vnl.end(), a, a + 3);
    } else if (c0 == 'u' && c1 == 's' && c2 == 'e') {
      use_pos.push_back(nface);
      use_name.push_back(after_last_space(l));
    } else if (l.compare(0, 2, "g ") == 0) {
      group_pos.push_back(nface);
      group_name.push_back(after_last_space(l));
    }
  }
  ocr_obj_mesh* m = new (std::nothrow) ocr_obj_mesh();
  if (!m) return OCR_ERR_OOM;
  const int64_t nv = (int64_t)vl.size() / 3, nvt = (int64_t)vtl.size() / 2, nvn = (int64_t)vnl.size() / 3;
  m->tris.resize(fl.size());
  for (size_t i = 0; i < fl.size(); i++) {
    ocr_triangle& t = m->tris[i];
    std::memset(&t, 0, sizeof(t));
    for (int j = 0; j < 3; j++) {
      const Idx& f = fl[i];
      if (f.v[j] < 0 || f.v[j] >= nv || f.vt[j] < 0 || f.vt[j] >= nvt || f.vn[j] < 0 || f.vn[j] >= nvn) {
        delete m;
        return OCR_ERR_INVALID;
      }
      ocr_vertex& v = t.vertices[j];
      for (int c = 0; c < 3; c++) v.pos[c] = vl[3 * f.v[j] + c] * scale;   // mult(v, requested_scale)
      v.pos[3] = 0.f * scale;
      v.vt[0] = vtl[2 * f.vt[j]];
      v.vt[1] = vtl[2 * f.vt[j] + 1];
      for (int c = 0; c < 3; c++) v.normal[c] = vnl[3 * f.vn[j] + c];
    }
  }
  const bool using_groups = use_pos.size() == group_pos.size();
  use_pos.push_back((uint32_t)m->tris.size());
  for (size_t i = 0; i + 1 < use_pos.size(); i++) {
    ocr_obj_part p;
    std::memset(&p, 0, sizeof(p));
    p.first_tri = use_pos[i];
    p.ntri = use_pos[i + 1] - use_pos[i];
    const std::string tname = diffuse_name(mtl, use_name[i]);
    const std::string bname = bump_name(mtl, use_name[i]);
    std::string full = dir + "/" + tname;
    const std::string cache = full;
    bool exists = file_exists(full);
    if (!exists && !tname.empty()) {
      const size_t s = tname.find_last_of("/\\");
      full = dir + "/" + tname.substr(s == std::string::npos ? 0 : s + 1);   // get_file_name
      exists = file_exists(full);
    }
    if (!tname.empty() && exists) {
      p.texture_kind = OCR_OBJ_TEX_FILE;
      copy_str(p.texture_cache, sizeof(p.texture_cache), cache);
      copy_str(p.texture_path, sizeof(p.texture_path), full);
    } else {
      float kd[3];
      if (kd_of(mtl, use_name[i], kd)) {
        const float r = kd[0] * 255.f, g = kd[1] * 255.f, b = kd[2] * 255.f;
        p.texture_kind = OCR_OBJ_TEX_COLOUR;
        p.colour[0] = to_u8(r);
        p.colour[1] = to_u8(g);
        p.colour[2] = to_u8(b);
        p.colour[3] = 255;
        copy_str(p.texture_cache, sizeof(p.texture_cache), colour_key(r, g, b));
      } else {
        p.texture_kind = OCR_OBJ_TEX_MAGENTA;
        p.colour[0] = 255;
        p.colour[1] = 0;
        p.colour[2] = 255;
        p.colour[3] = 255;
      }
    }
    if (bname != "None") copy_str(p.bump_path, sizeof(p.bump_path), dir + "/" + bname);
    if (using_groups) copy_str(p.group_name, sizeof(p.group_name), group_name[i]);
    m->parts.push_back(p);
  }
  *out = m;
  return OCR_OK;
}

uint32_t ocr_obj_ntri(const ocr_obj_mesh* m) { return m ? (uint32_t)m->tris.size() : 0; }
uint32_t ocr_obj_nparts(const ocr_obj_mesh* m) { return m ? (uint32_t)m->parts.size() : 0; }

int ocr_obj_get(const ocr_obj_mesh* m, ocr_triangle* tris, ocr_obj_part* parts) {
  if (!m) return OCR_ERR_INVALID;
  if (tris && !m->tris.empty()) std::memcpy(tris, m->tris.data(), sizeof(ocr_triangle) * m->tris.size());
  if (parts && !m->parts.empty()) std::memcpy(parts, m->parts.data(), sizeof(ocr_obj_part) * m->parts.size());
  return OCR_OK;
}

void ocr_obj_free(ocr_obj_mesh* m) { delete m; }

int ocr_png_load(const char* path, uint32_t* width, uint32_t* height, uint8_t** rgba) {
  if (!path || !width || !height || !rgba) return OCR_ERR_INVALID;
  *rgba = nullptr;
  std::ifstream f(path, std::ios::binary);
  if (!f.is_open()) return OCR_ERR_INVALID;
  std::vector<uint8_t> d((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
  static const uint8_t sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
  if (d.size() < 8 || std::memcmp(d.data(), sig, 8) != 0) return OCR_ERR_INVALID;
  uint32_t w = 0, h = 0;
  int depth = 0, ctype = -1, interlace = 0;
  std::vector<uint8_t> idat, plte, trns;
  for (size_t p = 8; p + 12 <= d.size();) {
    const uint32_t len = be32(&d[p]);
    if (p + 12 + (size_t)len > d.size()) return OCR_ERR_INVALID;
    const char* type = (const char*)&d[p + 4];
    const uint8_t* body = &d[p + 8];
    if (!std::memcmp(type, "IHDR", 4) && len >= 13) {
      w = be32(body);
      h = be32(body + 4);
      depth = body[8];
      ctype = body[9];
      interlace = body[12];
    } else if (!std::memcmp(type, "PLTE", 4)) {
      plte.assign(body, body + len);
    } else if (!std::memcmp(type, "tRNS", 4)) {
      trns.assign(body, body + len);
    } else if (!std::memcmp(type, "IDAT", 4)) {
      idat.insert(idat.end(), body, body + len);
    } else if (!std::memcmp(type, "IEND", 4)) {
      break;
    }
    p += 12 + (size_t)len;
  }
  const int chans = ctype == 0 ? 1 : ctype == 2 ? 3 : ctype == 3 ? 1 : ctype == 4 ? 2 : ctype == 6 ? 4 : 0;
  if (!w || !h || w > 32768 || h > 32768 || depth != 8 || !chans || interlace != 0) return OCR_ERR_UNSUPPORTED;
  if (ctype == 3 && plte.size() < 3) return OCR_ERR_INVALID;
  const size_t stride = (size_t)w * chans;
  // deflate expands at most 1032:1, so a header promising more than that
  // is refused before its buffer is allocated
  if ((uint64_t)h * (stride + 1) > (uint64_t)idat.size() * 1032 + 64) return OCR_ERR_INVALID;
  std::vector<uint8_t> raw(h * (stride + 1));
  uLongf raw_len = (uLongf)raw.size();
  if (uncompress(raw.data(), &raw_len, idat.data(), (uLong)idat.size()) != Z_OK || raw_len != raw.size())
    return OCR_ERR_INVALID;
  // reverse the per-row filters (None, Sub, Up, Average, Paeth)
  std::vector<uint8_t> px(h * stride);
  for (uint32_t y = 0; y < h; y++) {
    const uint8_t ft = raw[y * (stride + 1)];
    const uint8_t* src = &raw[y * (stride + 1) + 1];
    uint8_t* row = &px[y * stride];
    const uint8_t* up = y ? &px[(y - 1) * stride] : nullptr;
    for (size_t i = 0; i < stride; i++) {
      const int a = i >= (size_t)chans ? row[i - chans] : 0, b = up ? up[i] : 0,
                c = (up && i >= (size_t)chans) ? up[i - chans] : 0;
      int v;
      switch (ft) {
        case 0: v = src[i]; break;
        case 1: v = src[i] + a; break;
        case 2: v = src[i] + b; break;
        case 3: v = src[i] + ((a + b) >> 1); break;
        case 4: v = src[i] + paeth(a, b, c); break;
        default: return OCR_ERR_INVALID;
      }
      row[i] = (uint8_t)v;
    }
  }
  uint8_t* o = (uint8_t*)std::malloc((size_t)w * h * 4);
  if (!o) return OCR_ERR_OOM;
  for (size_t i = 0; i < (size_t)w * h; i++) {
    const uint8_t* s = &px[i * chans];
    uint8_t* t = &o[i * 4];
    switch (ctype) {
      case 0: t[0] = t[1] = t[2] = s[0]; t[3] = 255; break;
      case 4: t[0] = t[1] = t[2] = s[0]; t[3] = s[1]; break;
      case 2: t[0] = s[0]; t[1] = s[1]; t[2] = s[2]; t[3] = 255; break;
      case 6: std::memcpy(t, s, 4); break;
      case 3: {
        const size_t k = s[0];
        if (3 * k + 2 >= plte.size()) {
          std::free(o);
          return OCR_ERR_INVALID;
        }
        t[0] = plte[3 * k];
        t[1] = plte[3 * k + 1];
        t[2] = plte[3 * k + 2];
        t[3] = k < trns.size() ? trns[k] : 255;
        break;
      }
    }
  }
  *width = w;
  *height = h;
  *rgba = o;
  return OCR_OK;
}

void ocr_png_free(uint8_t* rgba) { std::free(rgba); }

}  // extern "C"
