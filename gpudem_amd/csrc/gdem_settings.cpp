// demSettings.json loader: the settings the reference's main.cu reads
// (main.cu:57-135, 227), the particle material (ParticleSystem.cu:189-352),
// the time step rule and the inlet.  Host code only.
//
// Object keys match case-insensitively and numbers are read with strtod, as
// cJSON_GetObjectItem and cJSON's parser do.  Differences from the reference,
// all on inputs it rejects or mishandles:
//  * a missing "YoungModulus" with a legacy "k" (the stiffness itself, as in
//    the reference's own case files) is accepted instead of aborting;
//  * a missing "bruteOn" or "gpuOn" means 0 instead of aborting;
//  * the "file" inlet accepts "file": "path" as well as "file": {"name":
//    "path"} (the reference's get_json_string only accepts the latter form
//    while its sample settings use the former).
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/gdem.h"

namespace gdem {
namespace {

struct Json {
  enum Kind { NUL, BOOL, NUM, STR, ARR, OBJ } kind = NUL;
  double num = 0;
  bool b = false;
  std::string str;
  std::vector<Json> arr;
  std::vector<std::pair<std::string, Json>> obj;

  const Json* get(const char* key) const {
    if (kind != OBJ) return nullptr;
    for (const auto& kv : obj) {
      const std::string& k = kv.first;
      if (k.size() != std::strlen(key)) continue;
      bool same = true;
      for (size_t i = 0; i < k.size() && same; ++i)
        same = std::tolower((unsigned char)k[i]) == std::tolower((unsigned char)key[i]);
      if (same) return &kv.second;
    }
    return nullptr;
  }
};

class Parser {
 public:
  explicit Parser(const std::string& s) : s_(s) {}
  bool parse(Json* out) {
    ws();
    if (!value(out)) return false;
    ws();
    return i_ == s_.size();
  }

 private:
  const std::string& s_;
  size_t i_ = 0;
  void ws() {
    while (i_ < s_.size() && std::isspace((unsigned char)s_[i_])) ++i_;
  }
  bool lit(const char* w) {
    const size_t n = std::strlen(w);
    if (s_.compare(i_, n, w) != 0) return false;
    i_ += n;
    return true;
  }
  bool string(std::string* out) {
    if (i_ >= s_.size() || s_[i_] != '"') return false;
    ++i_;
    out->clear();
    while (i_ < s_.size() && s_[i_] != '"') {
      char c = s_[i_++];
      if (c == '\\') {
        if (i_ >= s_.size()) return false;
        const char e = s_[i_++];
        switch (e) {
          case 'n': c = '\n'; break;
          case 't': c = '\t'; break;
          case 'r': c = '\r'; break;
          case 'b': c = '\b'; break;
          case 'f': c = '\f'; break;
          case 'u': {
            if (i_ + 4 > s_.size()) return false;
            const unsigned cp = (unsigned)std::strtoul(s_.substr(i_, 4).c_str(), nullptr, 16);
            i_ += 4;
            c = cp < 0x80 ? (char)cp : '?';
            break;
          }
          default: c = e;
        }
      }
      out->push_back(c);
    }
    if (i_ >= s_.size()) return false;
    ++i_;
    return true;
  }
  bool value(Json* v) {
    ws();
    if (i_ >= s_.size()) return false;
    const char c = s_[i_];
    if (c == '{') {
      ++i_;
      v->kind = Json::OBJ;
      ws();
      if (i_ < s_.size() && s_[i_] == '}') {
        ++i_;
        return true;
      }
      for (;;) {
        ws();
        std::string key;
        if (!string(&key)) return false;
        ws();
        if (i_ >= s_.size() || s_[i_] != ':') return false;
        ++i_;
        Json item;
        if (!value(&item)) return false;
        v->obj.emplace_back(key, std::move(item));
        ws();
        if (i_ < s_.size() && s_[i_] == ',') {
          ++i_;
          continue;
        }
        if (i_ < s_.size() && s_[i_] == '}') {
          ++i_;
          return true;
        }
        return false;
      }
    }
    if (c == '[') {
      ++i_;
      v->kind = Json::ARR;
      ws();
      if (i_ < s_.size() && s_[i_] == ']') {
        ++i_;
        return true;
      }
      for (;;) {
        Json item;
        if (!value(&item)) return false;
        v->arr.push_back(std::move(item));
        ws();
        if (i_ < s_.size() && s_[i_] == ',') {
          ++i_;
          continue;
        }
        if (i_ < s_.size() && s_[i_] == ']') {
          ++i_;
          return true;
        }
        return false;
      }
    }
    if (c == '"') {
      v->kind = Json::STR;
      return string(&v->str);
    }
    if (lit("true")) {
      v->kind = Json::BOOL;
      v->b = true;
      return true;
    }
    if (lit("false")) {
      v->kind = Json::BOOL;
      return true;
    }
    if (lit("null")) {
      v->kind = Json::NUL;
      return true;
    }
    // number: the longest prefix strtod accepts
    const char* start = s_.c_str() + i_;
    char* end = nullptr;
    v->num = std::strtod(start, &end);
    if (end == start) return false;
    v->kind = Json::NUM;
    i_ += (size_t)(end - start);
    return true;
  }
};

bool read_file(const char* path, std::string* out) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return false;
  std::ostringstream ss;
  ss << f.rdbuf();
  *out = ss.str();
  return true;
}

bool load_json(const char* path, Json* root) {
  std::string text;
  if (!path || !read_file(path, &text)) return false;
  return Parser(text).parse(root) && root->kind == Json::OBJ;
}

std::string dir_of(const char* path) {
  std::string p(path);
  const size_t k = p.find_last_of('/');
  return k == std::string::npos ? std::string(".") : p.substr(0, k);
}

std::string resolve(const char* settings_path, const std::string& rel) {
  if (!rel.empty() && rel[0] == '/') return rel;
  return dir_of(settings_path) + "/" + rel;
}

bool num(const Json* o, const char* key, double* out) {
  const Json* v = o ? o->get(key) : nullptr;
  if (!v || v->kind != Json::NUM) return false;
  *out = v->num;
  return true;
}

// cJSON's valueint: the double saturated to int range, then truncated
int as_int(double d) {
  if (d >= 2147483647.0) return 2147483647;
  if (d <= -2147483648.0) return (-2147483647 - 1);
  return (int)d;
}

// path of the particle file of a "file" inlet
bool inlet_file(const Json* inlet, const char* settings_path, std::string* out) {
  const Json* f = inlet->get("file");
  if (!f) return false;
  if (f->kind == Json::STR) {
    *out = resolve(settings_path, f->str);
    return true;
  }
  const Json* nm = f->get("name");
  if (!nm || nm->kind != Json::STR) return false;
  *out = resolve(settings_path, nm->str);
  return true;
}

// A CSV number as the reference's reader converts it (parse_float in csv.h,
// used by initializeParticles' file inlet): digits accumulated left to right
// (x = 10 x + d), fraction digits added as d * 10^-k with the power formed by
// repeated division, and an exponent applied by square-and-multiply powers of
// 10 or 0.1.  This is not correctly rounded, so strtod would differ in the
// last bits for some inputs.
double csv_number(const char* c) {
  bool neg = false;
  if (*c == '-') {
    neg = true;
    ++c;
  } else if (*c == '+') {
    ++c;
  }
  double x = 0;
  while ('0' <= *c && *c <= '9') {
    x *= 10;
    x += *c - '0';
    ++c;
  }
  if (*c == '.' || *c == ',') {
    ++c;
    double pos = 1;
    while ('0' <= *c && *c <= '9') {
      pos /= 10;
      x += (*c - '0') * pos;
      ++c;
    }
  }
  if (*c == 'e' || *c == 'E') {
    ++c;
    bool eneg = false;
    if (*c == '-') {
      eneg = true;
      ++c;
    } else if (*c == '+') {
      ++c;
    }
    long long e = 0;
    while ('0' <= *c && *c <= '9') {
      e = e * 10 + (*c - '0');
      if (e > 2147483647LL) e = 2147483647LL;
      ++c;
    }
    int ee = (int)e;
    if (ee != 0) {
      double base = eneg ? 0.1 : 10.;
      while (ee != 1) {
        if ((ee & 1) == 0) {
          base = base * base;
          ee >>= 1;
        } else {
          x *= base;
          --ee;
        }
      }
      x *= base;
    }
  }
  return neg ? -x : x;
}

// CSV with a header naming x, y, z, vx, vy, vz (any order, other columns
// ignored), one particle per row.
bool read_particle_csv(const std::string& path, std::vector<double>* x, std::vector<double>* v) {
  std::ifstream f(path);
  if (!f) return false;
  std::string line;
  if (!std::getline(f, line)) return false;
  auto split = [](const std::string& l) {
    std::vector<std::string> out;
    std::string cur;
    for (char c : l) {
      if (c == ',') {
        out.push_back(cur);
        cur.clear();
      } else if (c != '\r') {
        cur.push_back(c);
      }
    }
    out.push_back(cur);
    for (auto& s : out) {
      const size_t a = s.find_first_not_of(" \t"), b = s.find_last_not_of(" \t");
      s = a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
    }
    return out;
  };
  const std::vector<std::string> head = split(line);
  const char* names[6] = {"x", "y", "z", "vx", "vy", "vz"};
  int col[6];
  for (int k = 0; k < 6; ++k) {
    col[k] = -1;
    for (size_t c = 0; c < head.size(); ++c)
      if (head[c] == names[k]) col[k] = (int)c;
    if (col[k] < 0) return false;
  }
  while (std::getline(f, line)) {
    if (line.find_first_not_of(" \t\r") == std::string::npos) continue;
    const std::vector<std::string> f6 = split(line);
    double val[6];
    for (int k = 0; k < 6; ++k) {
      if ((size_t)col[k] >= f6.size()) return false;
      val[k] = csv_number(f6[col[k]].c_str());
    }
    x->insert(x->end(), val, val + 3);
    v->insert(v->end(), val + 3, val + 6);
  }
  return true;
}

}  // namespace
}  // namespace gdem

using namespace gdem;

extern "C" {

int gdem_settings_load(const char* path, gdem_settings* out) {
  if (!path || !out) return GDEM_E_ARG;
  std::memset(out, 0, sizeof(*out));
  Json root;
  if (!load_json(path, &root)) return GDEM_E_ARG;
  const Json* types = root.get("particleTypes");
  const Json* p0 = types ? types->get("particle0") : nullptr;
  const Json* inlet = root.get("inlet");
  const Json* others = root.get("others");
  const Json* grav = root.get("gravity");
  const Json* walls = root.get("walls");
  if (!p0 || !inlet || !others || !grav) return GDEM_E_ARG;
  double r, cor, density, mu, nparticles, dt_factor, out_time, end_time;
  if (!num(p0, "radius", &r) || !num(p0, "CoR", &cor) || !num(p0, "density", &density) || !num(p0, "mu", &mu))
    return GDEM_E_ARG;
  if (!num(inlet, "numParticle", &nparticles) || !num(others, "dtFactor", &dt_factor) ||
      !num(others, "outputTiming", &out_time) || !num(others, "endTime", &end_time))
    return GDEM_E_ARG;
  if (!num(grav, "x", &out->g[0]) || !num(grav, "y", &out->g[1]) || !num(grav, "z", &out->g[2]))
    return GDEM_E_ARG;
  if (!(r > 0) || !(density > 0) || !(cor > 0) || as_int(nparticles) <= 0) return GDEM_E_ARG;
  out->n = as_int(nparticles);
  out->radius = r;
  out->density = density;
  out->cor = cor;
  out->mu = mu;
  double young = 0, kk = 0;
  if (num(p0, "YoungModulus", &young)) {
    // ParticleSystem.cu:202
    kk = 4. / 3. * (young * 0.5) * sqrt(r * 0.5) * sqrt(0.05 * r);
    out->young = young;
  } else if (num(p0, "k", &kk)) {
    out->young = 0;
  } else {
    return GDEM_E_ARG;
  }
  out->k = kk;
  out->m = density * 3.14 * r * r * r * 4. / 3.;  // main.cu:67
  out->etaconst = -2. * log(cor) * sqrt(kk / (3.1415 * 3.1415 + log(cor) * log(cor)));  // ParticleSystem.cu:301
  // main.cu:123-129 and :227, PI from hardCodedParameters.h
  const double PI = 3.14159265359;
  double dt = 2. * PI * sqrt(out->m / kk) / dt_factor;
  const int out_step = (int)floor(out_time / dt);
  if (out_step <= 0) return GDEM_E_ARG;
  dt = out_time / (double)out_step;
  out->dt = dt;
  out->out_step = out_step;
  out->steps = (int)(end_time / dt);
  double v = 0;
  out->seed = num(others, "seed", &v) ? as_int(v) : 1;
  out->gpu_on = num(others, "gpuOn", &v) ? as_int(v) : 0;
  out->brute_on = num(others, "bruteOn", &v) ? as_int(v) : 0;
  const Json* fps = walls ? walls->get("filepaths") : nullptr;
  out->nwalls = (fps && fps->kind == Json::ARR) ? (int)fps->arr.size() : 0;
  // inlet box: the shuffle box, or the file's coordinate range
  // (calc_BoundingBoxLimits reads either, BoundingBox.cu:448-478)
  const Json* mode = inlet->get("inputMode");
  if (!mode || mode->kind != Json::STR) return GDEM_E_ARG;
  if (mode->str == "shuffle") {
    out->input_mode = GDEM_INLET_SHUFFLE;
    const Json* sh = inlet->get("shuffle");
    const char* keys[6] = {"minx", "maxx", "miny", "maxy", "minz", "maxz"};
    for (int k = 0; k < 6; ++k)
      if (!num(sh, keys[k], &out->inlet_box[k])) return GDEM_E_ARG;
  } else if (mode->str == "file" || mode->str == "files") {
    out->input_mode = GDEM_INLET_FILE;
    std::string file;
    std::vector<double> x, vv;
    if (!inlet_file(inlet, path, &file) || !read_particle_csv(file, &x, &vv)) return GDEM_E_ARG;
    if ((int)(x.size() / 3) != out->n) return GDEM_E_ARG;
    for (int a = 0; a < 3; ++a) {
      double lo = 1.e16, hi = -1.e16;
      for (int i = 0; i < out->n; ++i) {
        lo = x[3 * i + a] < lo ? x[3 * i + a] : lo;
        hi = x[3 * i + a] > hi ? x[3 * i + a] : hi;
      }
      out->inlet_box[2 * a] = lo;
      out->inlet_box[2 * a + 1] = hi;
    }
  } else {
    return GDEM_E_ARG;
  }
  return GDEM_OK;
}

int gdem_settings_wall(const char* path, int i, char* buf, int len) {
  if (!path || !buf || len <= 0) return GDEM_E_ARG;
  Json root;
  if (!load_json(path, &root)) return GDEM_E_ARG;
  const Json* walls = root.get("walls");
  const Json* fps = walls ? walls->get("filepaths") : nullptr;
  if (!fps || fps->kind != Json::ARR || i < 0 || i >= (int)fps->arr.size() || fps->arr[i].kind != Json::STR)
    return GDEM_E_ARG;
  const std::string full = resolve(path, fps->arr[i].str);
  if ((int)full.size() + 1 > len) return GDEM_E_ARG;
  std::memcpy(buf, full.c_str(), full.size() + 1);
  return GDEM_OK;
}

int gdem_settings_particles(const char* path, double* x, double* v) {
  if (!path || !x || !v) return GDEM_E_ARG;
  gdem_settings st;
  int rc = gdem_settings_load(path, &st);
  if (rc) return rc;
  const int n = st.n;
  if (st.input_mode == GDEM_INLET_FILE) {
    Json root;
    load_json(path, &root);
    std::string file;
    std::vector<double> xx, vv;
    if (!inlet_file(root.get("inlet"), path, &file) || !read_particle_csv(file, &xx, &vv)) return GDEM_E_ARG;
    std::memcpy(x, xx.data(), sizeof(double) * 3 * n);
    std::memcpy(v, vv.data(), sizeof(double) * 3 * n);
    return GDEM_OK;
  }
  // initializeParticles, "shuffle" (ParticleSystem.cu:189-352): uniform
  // draws in the inlet box from std::mt19937 seeded with 1, each rejected
  // while it lies within 2r of an already placed particle, at most 1000
  // trials per particle.
  const int max_trials = 1000;
  std::mt19937 mt(1);
  std::uniform_real_distribution<double> uni(0, 1);
  const double* b = st.inlet_box;
  const double r = st.radius;
  for (int i = 0; i < n; ++i) {
    int trial = 0;
    while (trial < max_trials) {
      const double px = uni(mt) * (b[1] - b[0]) + b[0];
      const double py = uni(mt) * (b[3] - b[2]) + b[2];
      const double pz = uni(mt) * (b[5] - b[4]) + b[4];
      bool overlap = false;
      for (int j = 0; j < i && !overlap; ++j) {
        const double dx = px - x[j * 3 + 0];
        const double dy = py - x[j * 3 + 1];
        const double dz = pz - x[j * 3 + 2];
        const double d2 = dx * dx + dy * dy + dz * dz;
        if (d2 < 4 * r * r) overlap = true;
      }
      if (!overlap) {
        x[i * 3 + 0] = px;
        x[i * 3 + 1] = py;
        x[i * 3 + 2] = pz;
        break;
      }
      ++trial;
    }
    if (trial == max_trials) return GDEM_E_ARG;  // the reference aborts here
    v[i * 3 + 0] = v[i * 3 + 1] = v[i * 3 + 2] = 0.;
  }
  return GDEM_OK;
}

}  // extern "C"
