// Scenario files in the reference simulator's formats, read into host arrays
// for nbg_sys_set_state / nbg_sys_accelerations_at.  Host code only.
//
// JSON (Simulation::from_json, src/NBody/Simulation/Simulation.cpp:154-228):
// an object whose "particles" array holds objects with optional "position",
// "velocity", "acceleration" ({"x", "y", "z"}), "mass", "passive", "color"
// and "sphere".  A particle with a mass has a Mass (active); it has an
// Acceleration (is moved and receives forces) when the file gives one or
// when "passive" is absent or true (Simulation.cpp:208-211).
//
// Tipsy (Simulation::from_tipsy, Simulation.cpp:273-395): an FP64 time and
// five u32 counts (n, ndim, ngas, ndark, nstar), optionally padded to 32
// bytes; FP32 records of 12 (gas), 9 (dark) and 11 (star) words, mass,
// position and velocity first.  Read in the other byte order when ndim is
// outside 1..3.  The stars are loaded, the gas only when there are no
// stars, the dark matter never; every loaded particle is active and passive.

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/nbg.h"
#include "host_util.hpp"

using nbg::fail;

struct nbg_scenario {
    std::vector<float> x, y, z, vx, vy, vz, m;
    std::vector<uint8_t> flags;
};

namespace {

bool read_file(const char *path, std::string &out) {
    FILE *f = std::fopen(path, "rb");
    if (!f) return false;
    char buf[1 << 16];
    size_t k;
    while ((k = std::fread(buf, 1, sizeof buf, f)) > 0) out.append(buf, k);
    const bool ok = !std::ferror(f);
    std::fclose(f);
    return ok;
}

// A small JSON reader: the values a scenario needs (objects, arrays,
// numbers, strings, true / false / null), held as a tree.
struct JValue {
    enum Kind { Null, Bool, Num, Str, Arr, Obj } kind = Null;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<JValue> items;                         // array elements, object values
    std::vector<std::string> keys;                     // object keys (parallel to items)
    const JValue *get(const char *k) const {
        if (kind != Obj) return nullptr;
        for (size_t i = 0; i < keys.size(); ++i)
            if (keys[i] == k) return &items[i];
        return nullptr;
    }
};

struct JParser {
    const char *p, *end;
    std::string err;
    int depth = 0;

    void ws() {
        while (p < end && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) ++p;
    }
    bool fail(const char *what) {
        if (err.empty()) err = what;
        return false;
    }
    bool lit(const char *s) {
        const size_t n = std::strlen(s);
        if ((size_t)(end - p) < n || std::strncmp(p, s, n) != 0) return false;
        p += n;
        return true;
    }
    bool string(std::string &out) {
        if (p >= end || *p != '"') return fail("expected a string");
        ++p;
        while (p < end && *p != '"') {
            if (*p == '\\') {
                if (++p >= end) return fail("unterminated string");
                const char c = *p++;
                switch (c) {
                case 'n': out += '\n'; break;
                case 't': out += '\t'; break;
                case 'r': out += '\r'; break;
                case 'b': out += '\b'; break;
                case 'f': out += '\f'; break;
                case 'u':  // keys and names here are ASCII: keep the escape as text
                    if (end - p < 4) return fail("bad \\u escape");
                    out += "\\u";
                    out.append(p, 4);
                    p += 4;
                    break;
                default: out += c;
                }
            } else {
                out += *p++;
            }
        }
        if (p >= end) return fail("unterminated string");
        ++p;
        return true;
    }
    bool value(JValue &v) {
        if (++depth > 64) return fail("nesting too deep");
        ws();
        bool ok = true;
        if (p >= end) {
            ok = fail("unexpected end of input");
        } else if (*p == '{') {
            v.kind = JValue::Obj;
            ++p;
            ws();
            if (p < end && *p == '}') {
                ++p;
            } else {
                while (ok) {
                    ws();
                    std::string k;
                    if (!(ok = string(k))) break;
                    ws();
                    if (p >= end || *p != ':') { ok = fail("expected ':'"); break; }
                    ++p;
                    v.keys.push_back(k);
                    v.items.emplace_back();
                    if (!(ok = value(v.items.back()))) break;
                    ws();
                    if (p < end && *p == ',') { ++p; continue; }
                    if (p < end && *p == '}') { ++p; break; }
                    ok = fail("expected ',' or '}'");
                }
            }
        } else if (*p == '[') {
            v.kind = JValue::Arr;
            ++p;
            ws();
            if (p < end && *p == ']') {
                ++p;
            } else {
                while (ok) {
                    v.items.emplace_back();
                    if (!(ok = value(v.items.back()))) break;
                    ws();
                    if (p < end && *p == ',') { ++p; continue; }
                    if (p < end && *p == ']') { ++p; break; }
                    ok = fail("expected ',' or ']'");
                }
            }
        } else if (*p == '"') {
            v.kind = JValue::Str;
            ok = string(v.str);
        } else if (lit("true")) {
            v.kind = JValue::Bool;
            v.b = true;
        } else if (lit("false")) {
            v.kind = JValue::Bool;
        } else if (lit("null")) {
            v.kind = JValue::Null;
        } else {
            char *q = nullptr;
            const std::string tok(p, (size_t)std::min<ptrdiff_t>(end - p, 64));
            v.num = std::strtod(tok.c_str(), &q);
            if (q == tok.c_str()) {
                ok = fail("bad value");
            } else {
                v.kind = JValue::Num;
                p += (q - tok.c_str());
            }
        }
        --depth;
        return ok;
    }
};

bool vec3(const JValue *v, float out[3]) {
    if (!v || v->kind != JValue::Obj) return false;
    const char *names[3] = {"x", "y", "z"};
    for (int d = 0; d < 3; ++d) {
        const JValue *c = v->get(names[d]);
        if (!c || c->kind != JValue::Num) return false;
        out[d] = (float)c->num;  // the reference reads each component as a float (glm::vec3)
    }
    return true;
}

uint32_t u32_at(const std::string &d, size_t off, bool swap) {
    uint32_t v;
    std::memcpy(&v, d.data() + off, 4);
    return swap ? __builtin_bswap32(v) : v;
}

float f32_at(const std::string &d, size_t off, bool swap) {
    const uint32_t u = u32_at(d, off, swap);
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

}  // namespace

extern "C" {

int nbg_scenario_load_json(const char *path, nbg_scenario **out) {
    if (!path || !out) return fail(NBG_ERR_ARG, "path or out is NULL");
    *out = nullptr;
    std::string text;
    if (!read_file(path, text)) return fail(NBG_ERR_ARG, std::string("cannot read ") + path);
    JParser P{text.data(), text.data() + text.size()};
    JValue root;
    if (!P.value(root)) return fail(NBG_ERR_ARG, std::string("JSON: ") + P.err);
    P.ws();
    if (P.p != P.end) return fail(NBG_ERR_ARG, "JSON: trailing characters");
    const JValue *parts = root.get("particles");
    if (!parts || parts->kind != JValue::Arr) return fail(NBG_ERR_ARG, "JSON scenario without a \"particles\" array");
    nbg_scenario *s = new nbg_scenario;
    const size_t n = parts->items.size();
    for (auto *v : {&s->x, &s->y, &s->z, &s->vx, &s->vy, &s->vz, &s->m}) v->assign(n, 0.0f);
    s->flags.assign(n, 0);
    for (size_t i = 0; i < n; ++i) {
        const JValue &e = parts->items[i];
        if (e.kind != JValue::Obj) {
            delete s;
            return fail(NBG_ERR_ARG, "JSON: a particle is not an object");
        }
        float v[3];
        uint8_t f = 0;
        if (const JValue *q = e.get("position")) {
            if (!vec3(q, v)) { delete s; return fail(NBG_ERR_ARG, "JSON: bad position"); }
            s->x[i] = v[0]; s->y[i] = v[1]; s->z[i] = v[2];
            f |= NBG_SCENARIO_POSITION;
        }
        if (const JValue *q = e.get("velocity")) {
            if (!vec3(q, v)) { delete s; return fail(NBG_ERR_ARG, "JSON: bad velocity"); }
            s->vx[i] = v[0]; s->vy[i] = v[1]; s->vz[i] = v[2];
            f |= NBG_SCENARIO_VELOCITY;
        }
        if (const JValue *q = e.get("mass")) {
            if (q->kind != JValue::Num) { delete s; return fail(NBG_ERR_ARG, "JSON: bad mass"); }
            s->m[i] = (float)q->num;
            f |= NBG_SCENARIO_ACTIVE;
        }
        const JValue *pa = e.get("passive");
        if (e.get("acceleration") || !pa || (pa->kind == JValue::Bool && pa->b)) f |= NBG_SCENARIO_PASSIVE;
        s->flags[i] = f;
    }
    *out = s;
    return NBG_OK;
}

int nbg_scenario_load_tipsy(const char *path, nbg_scenario **out) {
    if (!path || !out) return fail(NBG_ERR_ARG, "path or out is NULL");
    *out = nullptr;
    std::string d;
    if (!read_file(path, d)) return fail(NBG_ERR_ARG, std::string("cannot read ") + path);
    if (d.size() < 28) return fail(NBG_ERR_ARG, "tipsy: shorter than a header");
    bool swap = false;
    uint32_t ndim = u32_at(d, 12, false);
    if (ndim < 1 || ndim > 3) {
        swap = true;
        ndim = u32_at(d, 12, true);
        if (ndim < 1 || ndim > 3) return fail(NBG_ERR_ARG, "tipsy: no plausible dimension count in either byte order");
    }
    const uint64_t ng = u32_at(d, 16, swap), nd = u32_at(d, 20, swap), ns = u32_at(d, 24, swap);
    const uint64_t body = 48 * ng + 36 * nd + 44 * ns;
    size_t off;
    if (d.size() == 32 + body) off = 32;
    else if (d.size() == 28 + body) off = 28;
    else return fail(NBG_ERR_ARG, "tipsy: the file size does not match the header");
    const size_t gas_off = off, star_off = off + 48 * ng + 36 * nd;
    nbg_scenario *s = new nbg_scenario;
    auto take = [&](size_t base, uint64_t count, size_t stride) {
        for (uint64_t k = 0; k < count; ++k) {
            const size_t r = base + k * stride;
            s->m.push_back(f32_at(d, r, swap));
            s->x.push_back(f32_at(d, r + 4, swap));
            s->y.push_back(f32_at(d, r + 8, swap));
            s->z.push_back(f32_at(d, r + 12, swap));
            s->vx.push_back(f32_at(d, r + 16, swap));
            s->vy.push_back(f32_at(d, r + 20, swap));
            s->vz.push_back(f32_at(d, r + 24, swap));
            s->flags.push_back(NBG_SCENARIO_POSITION | NBG_SCENARIO_VELOCITY | NBG_SCENARIO_ACTIVE |
                               NBG_SCENARIO_PASSIVE);
        }
    };
    if (ns == 0) take(gas_off, ng, 48);
    take(star_off, ns, 44);
    *out = s;
    return NBG_OK;
}

int64_t nbg_scenario_count(const nbg_scenario *s) { return s ? (int64_t)s->m.size() : -1; }

int nbg_scenario_get(const nbg_scenario *s, float *x, float *y, float *z, float *vx, float *vy, float *vz, float *m,
                     uint8_t *flags) {
    if (!s) return fail(NBG_ERR_ARG, "scenario is NULL");
    const size_t n = s->m.size();
    float *dst[7] = {x, y, z, vx, vy, vz, m};
    const std::vector<float> *src[7] = {&s->x, &s->y, &s->z, &s->vx, &s->vy, &s->vz, &s->m};
    for (int k = 0; k < 7; ++k)
        if (dst[k] && n) std::memcpy(dst[k], src[k]->data(), n * sizeof(float));
    if (flags && n) std::memcpy(flags, s->flags.data(), n);
    return NBG_OK;
}

int nbg_scenario_free(nbg_scenario *s) {
    delete s;
    return NBG_OK;
}

}  // extern "C"
