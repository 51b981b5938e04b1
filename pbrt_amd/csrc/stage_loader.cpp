// The reference's scene files on the device: the `.stage` language (src/loader/parser/lexer.rs,
// src/loader/parser.rs), the scene builder's semantics (src/loader/visitors/scene_builder.rs) and
// the ASCII PLY meshes it loads (src/loader/ply/ply_loader.rs:251-255, src/loader/mesh_loader.rs),
// ending in a committed pbrt_scene with the lights Loader::load_scene adds (src/loader.rs:29-46).
//
// Builder semantics kept: the scene receives its objects in reverse file order (visit_scene pops
// them off a stack), a compound's children and a CSG's elements likewise; a transform block's
// steps apply in listed order (the identity times the steps from the last to the first); a mesh
// `reverse` flips every face's winding; vertex coordinates are read as f32 (`parse::<f32>`, here
// strtof, both correctly rounded). Differences: a compound object becomes its children as separate
// objects and nested transforms are composed into one (the same surfaces; the nested form
// renormalises the direction once per level, so hit distances can differ in the last bits);
// cylinders, planes and the thin-lens camera have no device counterpart and are refused.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "common.hpp"

namespace pbrt {
namespace {

// ---- lexer (lexer.rs) ----------------------------------------------------------------------------
struct Tok {
  enum Kind { kWord, kNumber, kString, kOpen, kClose, kEnd } kind = kEnd;
  std::string text;
  double num = 0.0;
  int line = 0;
};

class Lexer {
 public:
  explicit Lexer(std::string s) : s_(std::move(s)) {}
  Tok next() {
    if (has_back_) {
      has_back_ = false;
      return back_;
    }
    for (;;) {
      while (i_ < s_.size() && std::isspace((unsigned char)s_[i_])) {
        if (s_[i_] == '\n') ++line_;
        ++i_;
      }
      if (i_ < s_.size() && s_[i_] == '#') {  // a comment runs to the end of the line
        while (i_ < s_.size() && s_[i_] != '\n') ++i_;
        continue;
      }
      break;
    }
    Tok t;
    t.line = line_;
    if (i_ >= s_.size()) return t;
    const char c = s_[i_];
    if (c == '{' || c == '}') {
      t.kind = c == '{' ? Tok::kOpen : Tok::kClose;
      ++i_;
      return t;
    }
    if (c == '"') {
      const size_t e = s_.find('"', i_ + 1);
      if (e == std::string::npos) fail(t.line, "unterminated string");
      t.kind = Tok::kString;
      t.text = s_.substr(i_ + 1, e - i_ - 1);
      i_ = e + 1;
      return t;
    }
    if (std::isdigit((unsigned char)c) || c == '-' || c == '+') {
      // [sign] digits [. digits], no exponent (lexer.rs tokenize_number)
      size_t j = i_;
      if (s_[j] == '-' || s_[j] == '+') ++j;
      const size_t d0 = j;
      while (j < s_.size() && std::isdigit((unsigned char)s_[j])) ++j;
      if (j == d0) fail(t.line, "malformed number");
      if (j < s_.size() && s_[j] == '.') {
        ++j;
        while (j < s_.size() && std::isdigit((unsigned char)s_[j])) ++j;
      }
      t.kind = Tok::kNumber;
      t.text = s_.substr(i_, j - i_);
      t.num = std::strtod(t.text.c_str(), nullptr);
      i_ = j;
      return t;
    }
    if (std::isalpha((unsigned char)c)) {
      size_t j = i_;
      while (j < s_.size() && (std::isalnum((unsigned char)s_[j]) || s_[j] == '_')) ++j;
      t.kind = Tok::kWord;
      t.text = s_.substr(i_, j - i_);
      i_ = j;
      return t;
    }
    fail(t.line, std::string("unexpected character '") + c + "'");
    return t;
  }
  void take_back(const Tok& t) {
    back_ = t;
    has_back_ = true;
  }
  [[noreturn]] static void fail(int line, const std::string& what) {
    throw Error(PBRT_ERR_INVALID_ARG, ".stage line " + std::to_string(line) + ": " + what);
  }

 private:
  std::string s_;
  size_t i_ = 0;
  int line_ = 1;
  Tok back_;
  bool has_back_ = false;
};

// ---- transforms (geom/transform.rs, geom/matrix4.rs) ---------------------------------------------
struct M4 {
  double a[4][4];
};
M4 identity() {
  M4 m{};
  for (int i = 0; i < 4; ++i) m.a[i][i] = 1.0;
  return m;
}
M4 mat_mul(const M4& x, const M4& y) {  // matrix4.rs:216-232: each entry summed from 0 over k
  M4 r{};
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      double s = 0.0;
      for (int k = 0; k < 4; ++k) s += x.a[i][k] * y.a[k][j];
      r.a[i][j] = s;
    }
  return r;
}
struct Xf {
  M4 m = identity(), inv = identity();
  bool set = false;
};
Xf xf_mul(const Xf& x, const Xf& y) {  // transform.rs:119-127
  Xf r;
  r.m = mat_mul(x.m, y.m);
  r.inv = mat_mul(y.inv, x.inv);
  r.set = x.set || y.set;
  return r;
}
Xf translation(double x, double y, double z) {  // transform.rs:37-42
  Xf t;
  t.m.a[0][3] = x;
  t.m.a[1][3] = y;
  t.m.a[2][3] = z;
  t.inv.a[0][3] = -x;
  t.inv.a[1][3] = -y;
  t.inv.a[2][3] = -z;
  t.set = true;
  return t;
}
Xf rotation(char axis, double theta) {  // transform.rs:44-60, matrix4.rs:46-68
  const double s = std::sin(theta), c = std::cos(theta);
  Xf t;
  M4& m = t.m;
  if (axis == 'x') {
    m.a[1][1] = c; m.a[1][2] = -s; m.a[2][1] = s; m.a[2][2] = c;
  } else if (axis == 'y') {
    m.a[0][0] = c; m.a[0][2] = s; m.a[2][0] = -s; m.a[2][2] = c;
  } else {
    m.a[0][0] = c; m.a[0][1] = -s; m.a[1][0] = s; m.a[1][1] = c;
  }
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) t.inv.a[i][j] = m.a[j][i];
  t.set = true;
  return t;
}

// ---- AST -----------------------------------------------------------------------------------------
struct ShapeN {
  std::string kind;  // aabox, sphere, rectangle, mesh, csg
  double param[3] = {0, 0, 0};
  std::string file;
  bool reverse = false;
  int op = 0;
  std::vector<std::pair<ShapeN, Xf>> elems;  // CSG elements in file order
};
struct MatN {
  int kind = 0;
  double rgb[3] = {0, 0, 0}, rgb2[3] = {0, 0, 0};
  int tex = 0;
  double scale = 0.0, ior = 1.5, fuzz = 0.0;
};
struct ObjN {
  enum Kind { kSimple, kTransformed, kCompound } kind = kSimple;
  ShapeN shape;
  MatN mat;
  Xf xf;
  std::vector<ObjN> children;  // transformed: one; compound: in file order
};
// A geometry-less light of the `light` production (planned in the reference, IDEAS.md:25-27 and
// 180-199: a scene member beside the objects, positioned by a transform block; area lights stay a
// diffuse_light material on an object).
struct LightN {
  enum Kind { kPoint, kUniform, kBackground } kind = kPoint;
  double rgb[3] = {0, 0, 0}, rgb2[3] = {0, 0, 0};
  Xf xf;
};
struct StageN {
  double pos[3], look[3], up[3];
  std::vector<ObjN> objects;
  std::vector<LightN> lights;  // file order
};

// ---- parser (parser.rs) --------------------------------------------------------------------------
class Parser {
 public:
  explicit Parser(std::string s) : lx_(std::move(s)) {}
  StageN stage() {
    StageN st{};
    word("camera");
    const Tok cam = lx_.next();
    if (cam.kind != Tok::kWord || cam.text != "pin_hole")
      Lexer::fail(cam.line, cam.text == "thin_lens" ? "the thin-lens camera is not supported" : "camera pin_hole expected");
    word("pos");
    vec(st.pos);
    word("look");
    vec(st.look);
    word("up");
    vec(st.up);
    word("scene");
    for (;;) {
      const Tok t = lx_.next();
      if (t.kind == Tok::kEnd) break;
      if (t.kind == Tok::kWord && t.text == "light") {
        st.lights.push_back(light());
        continue;
      }
      if (t.kind != Tok::kWord || t.text != "object") Lexer::fail(t.line, "object or light expected");
      st.objects.push_back(object());
    }
    return st;
  }

 private:
  Lexer lx_;
  void word(const char* w) {
    const Tok t = lx_.next();
    if (t.kind != Tok::kWord || t.text != w) Lexer::fail(t.line, std::string("'") + w + "' expected");
  }
  double num() {
    const Tok t = lx_.next();
    if (t.kind != Tok::kNumber) Lexer::fail(t.line, "number expected");
    return t.num;
  }
  void vec(double* v) {
    for (int i = 0; i < 3; ++i) v[i] = num();
  }
  void open() {
    const Tok t = lx_.next();
    if (t.kind != Tok::kOpen) Lexer::fail(t.line, "'{' expected");
  }
  ObjN object() {
    const Tok t = lx_.next();
    ObjN o;
    if (t.kind == Tok::kWord && t.text == "simple") {
      o.kind = ObjN::kSimple;
      o.shape = shape();
      o.mat = material();
    } else if (t.kind == Tok::kWord && t.text == "transformed") {
      o.kind = ObjN::kTransformed;
      word("object");
      o.children.push_back(object());
      o.xf = transform();
    } else if (t.kind == Tok::kWord && t.text == "compound") {
      o.kind = ObjN::kCompound;
      open();
      for (;;) {
        const Tok u = lx_.next();
        if (u.kind == Tok::kClose) break;
        if (u.kind != Tok::kWord || u.text != "object") Lexer::fail(u.line, "'}' or object expected");
        o.children.push_back(object());
      }
    } else {
      Lexer::fail(t.line, "simple, transformed or compound expected");
    }
    return o;
  }
  // light point color r g b [transform {...}] | light uniform_infinite color r g b |
  // light background_infinite color r g b color r g b (the colours at y = -1 and y = +1)
  LightN light() {
    const Tok t = lx_.next();
    LightN l;
    if (t.kind != Tok::kWord) Lexer::fail(t.line, "light kind expected");
    if (t.text == "point") {
      l.kind = LightN::kPoint;
      word("color");
      vec(l.rgb);
      const Tok x = lx_.next();
      lx_.take_back(x);
      if (x.kind == Tok::kWord && x.text == "transform") l.xf = transform();
    } else if (t.text == "uniform_infinite") {
      l.kind = LightN::kUniform;
      word("color");
      vec(l.rgb);
    } else if (t.text == "background_infinite") {
      l.kind = LightN::kBackground;
      word("color");
      vec(l.rgb);
      word("color");
      vec(l.rgb2);
    } else if (t.text == "area" || t.text == "diffuse_light") {
      Lexer::fail(t.line, "an area light is a diffuse_light material on a rectangle object, not a light");
    } else {
      Lexer::fail(t.line, "point, uniform_infinite or background_infinite expected");
    }
    return l;
  }
  ShapeN shape() {
    const Tok t = lx_.next();
    ShapeN s;
    if (t.kind != Tok::kWord) Lexer::fail(t.line, "shape expected");
    s.kind = t.text;
    if (t.text == "aabox") {
      vec(s.param);
    } else if (t.text == "sphere") {
      s.param[0] = num();
    } else if (t.text == "rectangle") {
      s.param[0] = num();
      s.param[1] = num();
    } else if (t.text == "mesh") {
      word("file");
      const Tok f = lx_.next();
      if (f.kind != Tok::kString) Lexer::fail(f.line, "mesh file name expected");
      s.file = f.text;
      const Tok r = lx_.next();
      if (r.kind == Tok::kWord && r.text == "reverse")
        s.reverse = true;
      else
        lx_.take_back(r);
    } else if (t.text == "csg") {
      const Tok o = lx_.next();
      if (o.kind == Tok::kWord && o.text == "union") s.op = PBRT_CSG_UNION;
      else if (o.kind == Tok::kWord && o.text == "intersection") s.op = PBRT_CSG_INTERSECTION;
      else if (o.kind == Tok::kWord && (o.text == "substraction" || o.text == "subtraction")) s.op = PBRT_CSG_SUBTRACTION;
      else Lexer::fail(o.line, "intersection, union or substraction expected");
      open();
      for (;;) {
        const Tok e = lx_.next();
        if (e.kind == Tok::kClose) break;
        if (e.kind != Tok::kWord || e.text != "elem") Lexer::fail(e.line, "'}' or elem expected");
        ShapeN el = shape();
        if (el.kind != "sphere" && el.kind != "aabox") Lexer::fail(e.line, "CSG elements are spheres and boxes");
        const Xf x = transform();
        s.elems.emplace_back(el, x);
      }
    } else if (t.text == "cylinder" || t.text == "plane") {
      Lexer::fail(t.line, "the " + t.text + " shape is not supported");
    } else {
      Lexer::fail(t.line, "unknown shape '" + t.text + "'");
    }
    return s;
  }
  MatN material() {
    const Tok t = lx_.next();
    MatN m;
    if (t.kind != Tok::kWord) Lexer::fail(t.line, "material expected");
    if (t.text == "lambertian") {
      m.kind = PBRT_MAT_LAMBERTIAN;
    } else if (t.text == "dielectric") {
      m.kind = PBRT_MAT_DIELECTRIC;
      m.ior = num();
    } else if (t.text == "diffuse_light") {
      m.kind = PBRT_MAT_DIFFUSE_LIGHT;
    } else if (t.text == "metal") {
      m.kind = PBRT_MAT_METAL;
      m.fuzz = num();
    } else {
      Lexer::fail(t.line, "lambertian, dielectric, diffuse_light or metal expected");
    }
    const Tok x = lx_.next();
    if (x.kind == Tok::kWord && x.text == "color") {
      vec(m.rgb);
    } else if (x.kind == Tok::kWord && x.text == "checkerboard") {
      m.tex = PBRT_TEX_CHECKER;
      vec(m.rgb);
      vec(m.rgb2);
      m.scale = num();
    } else {
      Lexer::fail(x.line, "color or checkerboard expected");
    }
    return m;
  }
  Xf transform() {
    word("transform");
    open();
    std::vector<Xf> steps;
    for (;;) {
      const Tok t = lx_.next();
      if (t.kind == Tok::kClose) break;
      if (t.kind != Tok::kWord) Lexer::fail(t.line, "transform step expected");
      if (t.text == "translate") {
        double v[3];
        vec(v);
        steps.push_back(translation(v[0], v[1], v[2]));
      } else if (t.text == "rotate_x" || t.text == "rotate_y" || t.text == "rotate_z") {
        steps.push_back(rotation(t.text.back(), num()));
      } else {
        Lexer::fail(t.line, "translate or rotate_x/y/z expected");
      }
    }
    // scene_builder.rs:195-207: the identity times the steps taken from the last to the first
    Xf acc;
    for (auto it = steps.rbegin(); it != steps.rend(); ++it) acc = xf_mul(acc, *it);
    acc.set = true;
    return acc;
  }
};

// ---- PLY (ply_loader.rs, ASCII only) -------------------------------------------------------------
struct Ply {
  std::vector<float> xyz;
  std::vector<uint32_t> idx;
};

Ply read_ply(const std::string& path, bool reverse) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw Error(PBRT_ERR_INVALID_ARG, "cannot open PLY file " + path);
  std::string line;
  auto bad = [&](const std::string& what) { throw Error(PBRT_ERR_INVALID_ARG, path + ": " + what); };
  if (!std::getline(f, line) || line.substr(0, 3) != "ply") bad("not a PLY file");
  struct Elem {
    std::string name;
    uint64_t count;
    std::vector<std::string> props;  // property names; "list" for a list property
  };
  std::vector<Elem> elems;
  for (;;) {
    if (!std::getline(f, line)) bad("header without end_header");
    std::istringstream ls(line);
    std::string w;
    ls >> w;
    if (w == "format") {
      std::string fmt;
      ls >> fmt;
      if (fmt != "ascii") bad("only ASCII PLY is read (as the reference's loader)");
    } else if (w == "element") {
      Elem e;
      ls >> e.name >> e.count;
      elems.push_back(e);
    } else if (w == "property") {
      if (elems.empty()) bad("property before any element");
      std::string type, name;
      ls >> type;
      if (type == "list") {
        std::string ct, it;
        ls >> ct >> it >> name;
        elems.back().props.push_back("list");
      } else {
        ls >> name;
        elems.back().props.push_back(name);
      }
    } else if (w == "end_header") {
      break;
    }
  }
  Ply out;
  for (const Elem& e : elems) {
    if (e.name == "vertex") {
      int ix[3] = {-1, -1, -1};
      for (size_t k = 0; k < e.props.size(); ++k) {
        if (e.props[k] == "x") ix[0] = (int)k;
        if (e.props[k] == "y") ix[1] = (int)k;
        if (e.props[k] == "z") ix[2] = (int)k;
      }
      if (ix[0] < 0 || ix[1] < 0 || ix[2] < 0) bad("vertex without x, y, z");
      out.xyz.reserve(3 * e.count);
      std::vector<std::string> tok(e.props.size());
      for (uint64_t i = 0; i < e.count; ++i) {
        if (!std::getline(f, line)) bad("truncated vertex list");
        std::istringstream ls(line);
        for (auto& s : tok)
          if (!(ls >> s)) bad("short vertex line");
        for (int c = 0; c < 3; ++c) out.xyz.push_back(std::strtof(tok[ix[c]].c_str(), nullptr));
      }
    } else if (e.name == "face") {
      if (e.props.size() != 1 || e.props[0] != "list") bad("a face element must hold one index list");
      out.idx.reserve(3 * e.count);
      for (uint64_t i = 0; i < e.count; ++i) {
        if (!std::getline(f, line)) bad("truncated face list");
        std::istringstream ls(line);
        long n = 0, a = 0, b = 0, c = 0;
        if (!(ls >> n) || n != 3) bad("only triangle faces are supported");
        if (!(ls >> a >> b >> c) || a < 0 || b < 0 || c < 0) bad("malformed face");
        if (reverse) std::swap(a, c);  // mesh_loader.rs: the face's indices in reverse order
        out.idx.push_back((uint32_t)a);
        out.idx.push_back((uint32_t)b);
        out.idx.push_back((uint32_t)c);
      }
    } else {
      for (uint64_t i = 0; i < e.count; ++i)
        if (!std::getline(f, line)) bad("truncated element " + e.name);
    }
  }
  const uint64_t nv = out.xyz.size() / 3;
  for (uint32_t i : out.idx)
    if (i >= nv) bad("face index out of range");
  return out;
}

// ---- building (scene_builder.rs) -----------------------------------------------------------------
struct Flat {
  const ObjN* obj;
  Xf xf;
};

// The scene's objects in the order the builder adds them: the list reversed (a stack), a
// compound's children likewise, transforms composed outside-in.
void flatten(const ObjN& o, const Xf& outer, std::vector<Flat>* out) {
  switch (o.kind) {
    case ObjN::kSimple: out->push_back({&o, outer}); break;
    case ObjN::kTransformed: flatten(o.children[0], outer.set ? xf_mul(outer, o.xf) : o.xf, out); break;
    case ObjN::kCompound:
      for (auto it = o.children.rbegin(); it != o.children.rend(); ++it) flatten(*it, outer, out);
      break;
  }
}

void check(int status) {
  if (status != PBRT_OK) throw Error(status, pbrt_last_error());
}

std::string join(const char* base, const std::string& rel) {
  if (!base || !*base || (!rel.empty() && rel[0] == '/')) return rel;
  std::string b(base);
  if (b.back() != '/') b += '/';
  return b + rel;
}

void flat16(const M4& m, double* out) {
  for (int i = 0; i < 16; ++i) out[i] = m.a[i / 4][i % 4];
}

}  // namespace
}  // namespace pbrt

using namespace pbrt;

int pbrt_ply_read(const char* path, int reverse, float* xyz_out, uint32_t* nv, uint32_t* idx_out, uint32_t* nt) {
  return guarded([&] {
    if (!path || !nv || !nt) throw Error(PBRT_ERR_INVALID_ARG, "null argument");
    const Ply p = read_ply(path, reverse != 0);
    const uint32_t v = (uint32_t)(p.xyz.size() / 3), t = (uint32_t)(p.idx.size() / 3);
    if (xyz_out && idx_out) {
      if (*nv < v || *nt < t) throw Error(PBRT_ERR_INVALID_ARG, "output arrays too small");
      std::memcpy(xyz_out, p.xyz.data(), p.xyz.size() * sizeof(float));
      std::memcpy(idx_out, p.idx.data(), p.idx.size() * sizeof(uint32_t));
    }
    *nv = v;
    *nt = t;
  });
}

int pbrt_scene_load_stage(pbrt_ctx* ctx, const char* path, const char* base_dir, const pbrt_bvh_params* mesh_params,
                          pbrt_scene** scene_out, pbrt_render_config* view_out) {
  pbrt_scene* s = nullptr;
  const int rc = guarded([&] {
    if (!path || !scene_out) throw Error(PBRT_ERR_INVALID_ARG, "null argument");
    std::ifstream f(path, std::ios::binary);
    if (!f) throw Error(PBRT_ERR_INVALID_ARG, std::string("cannot open ") + path);
    std::stringstream ss;
    ss << f.rdbuf();
    const StageN st = Parser(ss.str()).stage();

    std::vector<Flat> flat;
    for (auto it = st.objects.rbegin(); it != st.objects.rend(); ++it) flatten(*it, Xf{}, &flat);

    // every mesh file read before any device work, so a file error costs no allocation
    std::map<std::pair<std::string, bool>, Ply> plys;
    for (const Flat& fl : flat) {
      const ShapeN& sh = fl.obj->shape;
      const auto key = std::make_pair(sh.file, sh.reverse);
      if (sh.kind == "mesh" && !plys.count(key)) plys.emplace(key, read_ply(join(base_dir, sh.file), sh.reverse));
    }
    if (!ctx) throw Error(PBRT_ERR_INVALID_ARG, "null context");
    check(pbrt_scene_create(ctx, mesh_params, &s));
    std::map<std::pair<std::string, bool>, uint32_t> meshes;
    for (const Flat& fl : flat) {
      const ObjN& o = *fl.obj;
      uint32_t mat = 0;
      check(pbrt_scene_add_material(s, o.mat.kind, o.mat.rgb, o.mat.ior, o.mat.kind == PBRT_MAT_METAL ? o.mat.fuzz : 0.0,
                                    &mat));
      if (o.mat.tex) check(pbrt_scene_set_texture(s, mat, o.mat.tex, o.mat.rgb2, o.mat.scale));
      double m16[16], mi16[16];
      flat16(fl.xf.m, m16);
      flat16(fl.xf.inv, mi16);
      const double* M = fl.xf.set ? m16 : nullptr;
      const double* MI = fl.xf.set ? mi16 : nullptr;
      uint32_t id = 0;
      const ShapeN& sh = o.shape;
      if (sh.kind == "mesh") {
        const auto key = std::make_pair(sh.file, sh.reverse);
        auto found = meshes.find(key);
        if (found == meshes.end()) {
          const Ply& p = plys.at(key);
          uint32_t mid = 0;
          check(pbrt_scene_add_mesh(s, p.xyz.data(), (uint32_t)(p.xyz.size() / 3), p.idx.data(),
                                    (uint32_t)(p.idx.size() / 3), &mid));
          found = meshes.emplace(key, mid).first;
        }
        check(pbrt_scene_add_object(s, PBRT_SHAPE_MESH, (int)found->second, sh.param, mat, M, MI, &id));
      } else if (sh.kind == "csg") {
        // the builder pops the elements off a stack: the constructor receives them reversed
        const size_t n = sh.elems.size();
        std::vector<int> shapes(n);
        std::vector<double> prm(3 * n), em(16 * n), emi(16 * n);
        for (size_t k = 0; k < n; ++k) {
          const auto& e = sh.elems[n - 1 - k];
          shapes[k] = e.first.kind == "sphere" ? PBRT_SHAPE_SPHERE : PBRT_SHAPE_BOX;
          for (int c = 0; c < 3; ++c) prm[3 * k + c] = e.first.param[c];
          flat16(e.second.m, &em[16 * k]);
          flat16(e.second.inv, &emi[16 * k]);
        }
        check(pbrt_scene_add_csg(s, sh.op, (uint32_t)n, shapes.data(), prm.data(), em.data(), emi.data(), mat, M, MI,
                                 &id));
      } else {
        const int kind = sh.kind == "sphere" ? PBRT_SHAPE_SPHERE : sh.kind == "aabox" ? PBRT_SHAPE_BOX : PBRT_SHAPE_RECT;
        check(pbrt_scene_add_object(s, kind, -1, sh.param, mat, M, MI, &id));
        // a file that declares its lights also gets its emissive rectangles as area lights
        // (ideas/area_light.md section 4: the builder registers them as it assembles the object)
        if (!st.lights.empty() && kind == PBRT_SHAPE_RECT && o.mat.kind == PBRT_MAT_DIFFUSE_LIGHT && !o.mat.tex)
          check(pbrt_scene_add_area_light(s, id, o.mat.rgb));
      }
    }
    if (st.lights.empty()) {
      // Loader::load_scene: a point light of intensity 15 at (0, 2, 1) and the white-to-sky
      // background light
      const double lp[3] = {0.0, 2.0, 1.0}, li[3] = {15.0, 15.0, 15.0};
      const double down[3] = {1.0, 1.0, 1.0}, up[3] = {0.5, 0.7, 1.0};
      check(pbrt_scene_add_point_light(s, lp, li));
      check(pbrt_scene_add_background_light(s, down, up));
    }
    for (const LightN& l : st.lights) {
      if (l.kind == LightN::kPoint) {
        // PointLight::sample_li: the transform applied to the origin (lights/point_light.rs:24)
        const double p[3] = {l.xf.m.a[0][3], l.xf.m.a[1][3], l.xf.m.a[2][3]};
        check(pbrt_scene_add_point_light(s, p, l.rgb));
      } else if (l.kind == LightN::kUniform) {
        // UniformInfiniteLight: the background light with one colour (the same sphere sampling)
        check(pbrt_scene_add_background_light(s, l.rgb, l.rgb));
      } else {
        check(pbrt_scene_add_background_light(s, l.rgb, l.rgb2));
      }
    }
    check(pbrt_scene_commit(s));
    if (view_out) {
      for (int c = 0; c < 3; ++c) {
        view_out->eye[c] = st.pos[c];
        view_out->look[c] = st.look[c];
        view_out->up[c] = st.up[c];
      }
      // the renderer's defaults (config.rs:112-125): the file holds no field of view or clipping
      view_out->fov = 90.0 * M_PI / 180.0;
      view_out->near = 1e-4;
      view_out->far = 1000.0;
    }
    *scene_out = s;
  });
  if (rc != PBRT_OK && s) pbrt_scene_destroy(s);
  return rc;
}
