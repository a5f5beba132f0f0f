// gps_camera.cpp — camera-path files and the frames they define.
//
// The reference keeps camera keyframes in JSON (CameraLoader::CameraState,
// src/utils/camera_loader.cuh:26-37; the files under assets/cameras): per camera an
// id, image name, width, height, position, a 3x3 rotation (camera-to-world,
// rows as written), fy and fx.  A frame of a camera state is built the way
// the viewer builds it (fromCameraState, src/utils/camera_loader.cu:63-90;
// Transform, src/world/transform.cpp:21-44; main.cu:146-158):
//   basis  = transpose(rotation as glm reads it) (= the rows of the file)
//   model  = translate(position) * mat4_cast(quat_cast(basis))
//   view   = scale(1, -1, -1) * affineInverse(model)
//   proj   = perspective(fovy, width / height, near, far)
// with fovy = focal2fovdeg(fy, height) (src/world/camera.h:44-47) unless the
// caller overrides it.  Path playback (CameraPath, camera.h:53-106) moves
// through the keyframes at 1/24 s each: position by mix, rotation by slerp.
//
// The JSON reader accepts the subset these files use (objects, arrays,
// numbers, strings, true / false / null) and ignores unknown keys.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/gps.h"
#include "gps_hostmath.h"

using namespace gps;

namespace {

struct JVal {
    enum Kind { NUL, BOOL, NUM, STR, ARR, OBJ } kind = NUL;
    double num = 0.0;
    std::string str;
    std::vector<JVal> arr;
    std::vector<std::pair<std::string, JVal>> obj;
    const JVal* get(const char* key) const {
        for (const auto& kv : obj)
            if (kv.first == key) return &kv.second;
        return nullptr;
    }
};

struct JParser {
    const char* p;
    const char* end;
    bool ok = true;
    void ws() {
        while (p < end && (*p == ' ' || *p == '\n' || *p == '\r' || *p == '\t')) ++p;
    }
    bool lit(const char* s) {
        const size_t n = std::strlen(s);
        if ((size_t)(end - p) >= n && std::strncmp(p, s, n) == 0) {
            p += n;
            return true;
        }
        return false;
    }
    std::string string() {
        std::string out;
        ++p;  // opening quote
        while (p < end && *p != '"') {
            if (*p == '\\' && p + 1 < end) {
                ++p;
                switch (*p) {
                    case 'n': out += '\n'; break;
                    case 't': out += '\t'; break;
                    case 'r': out += '\r'; break;
                    case 'b': out += '\b'; break;
                    case 'f': out += '\f'; break;
                    case 'u': {  // \uXXXX: basic plane, encoded as UTF-8
                        if (end - p < 5) {
                            ok = false;
                            return out;
                        }
                        const unsigned cp = (unsigned)std::strtoul(std::string(p + 1, p + 5).c_str(), nullptr, 16);
                        if (cp < 0x80) {
                            out += (char)cp;
                        } else if (cp < 0x800) {
                            out += (char)(0xC0 | (cp >> 6));
                            out += (char)(0x80 | (cp & 0x3F));
                        } else {
                            out += (char)(0xE0 | (cp >> 12));
                            out += (char)(0x80 | ((cp >> 6) & 0x3F));
                            out += (char)(0x80 | (cp & 0x3F));
                        }
                        p += 4;
                        break;
                    }
                    default: out += *p; break;  // \" \\ \/
                }
                ++p;
            } else {
                out += *p++;
            }
        }
        if (p >= end) ok = false;
        else ++p;  // closing quote
        return out;
    }
    JVal value(int depth = 0) {
        JVal v;
        ws();
        if (p >= end || depth > 64) {
            ok = false;
            return v;
        }
        if (*p == '{') {
            v.kind = JVal::OBJ;
            ++p;
            ws();
            if (p < end && *p == '}') {
                ++p;
                return v;
            }
            while (ok) {
                ws();
                if (p >= end || *p != '"') {
                    ok = false;
                    break;
                }
                std::string key = string();
                ws();
                if (p >= end || *p != ':') {
                    ok = false;
                    break;
                }
                ++p;
                JVal item = value(depth + 1);
                v.obj.emplace_back(std::move(key), std::move(item));
                ws();
                if (p < end && *p == ',') {
                    ++p;
                    continue;
                }
                if (p < end && *p == '}') {
                    ++p;
                    break;
                }
                ok = false;
            }
        } else if (*p == '[') {
            v.kind = JVal::ARR;
            ++p;
            ws();
            if (p < end && *p == ']') {
                ++p;
                return v;
            }
            while (ok) {
                v.arr.push_back(value(depth + 1));
                ws();
                if (p < end && *p == ',') {
                    ++p;
                    continue;
                }
                if (p < end && *p == ']') {
                    ++p;
                    break;
                }
                ok = false;
            }
        } else if (*p == '"') {
            v.kind = JVal::STR;
            v.str = string();
        } else if (lit("true")) {
            v.kind = JVal::BOOL;
            v.num = 1.0;
        } else if (lit("false")) {
            v.kind = JVal::BOOL;
        } else if (lit("null")) {
            v.kind = JVal::NUL;
        } else {
            char* e = nullptr;
            const std::string tok(p, (size_t)std::min<ptrdiff_t>(end - p, 64));
            v.num = std::strtod(tok.c_str(), &e);
            if (e == tok.c_str()) ok = false;
            v.kind = JVal::NUM;
            p += (e - tok.c_str());
        }
        return v;
    }
};

bool num(const JVal* v, double* out) {
    if (!v || v->kind != JVal::NUM) return false;
    *out = v->num;
    return true;
}
bool vec(const JVal* v, float* out, int n) {
    if (!v || v->kind != JVal::ARR || (int)v->arr.size() != n) return false;
    for (int i = 0; i < n; ++i) {
        double d;
        if (!num(&v->arr[i], &d)) return false;
        out[i] = (float)d;
    }
    return true;
}

bool to_state(const JVal& o, gps_camera_state* s) {
    if (o.kind != JVal::OBJ) return false;
    std::memset(s, 0, sizeof(*s));
    double id, w, h, fy, fx;
    if (!num(o.get("id"), &id) || !num(o.get("width"), &w) || !num(o.get("height"), &h) ||
        !num(o.get("fy"), &fy) || !num(o.get("fx"), &fx))
        return false;
    const JVal* name = o.get("img_name");
    if (!name || name->kind != JVal::STR) return false;
    if (!vec(o.get("position"), s->position, 3)) return false;
    const JVal* rot = o.get("rotation");
    if (!rot || rot->kind != JVal::ARR || rot->arr.size() != 3) return false;
    for (int r = 0; r < 3; ++r)
        if (!vec(&rot->arr[r], s->rotation + 3 * r, 3)) return false;
    s->id = (int32_t)id;
    s->width = (int32_t)w;
    s->height = (int32_t)h;
    s->fy = (float)fy;
    s->fx = (float)fx;
    std::snprintf(s->img_name, sizeof(s->img_name), "%s", name->str.c_str());
    return true;
}

struct Quat {
    float w, x, y, z;
};

// glm::quat_cast of a 3x3 (gtc/quaternion.inl:81-123), m[c][r].
Quat quat_cast(const float m[3][3]) {
    const float fx = m[0][0] - m[1][1] - m[2][2];
    const float fy = m[1][1] - m[0][0] - m[2][2];
    const float fz = m[2][2] - m[0][0] - m[1][1];
    const float fw = m[0][0] + m[1][1] + m[2][2];
    int big = 0;
    float fb = fw;
    if (fx > fb) {
        fb = fx;
        big = 1;
    }
    if (fy > fb) {
        fb = fy;
        big = 2;
    }
    if (fz > fb) {
        fb = fz;
        big = 3;
    }
    const float bv = std::sqrt(fb + 1.0f) * 0.5f;
    const float mult = 0.25f / bv;
    switch (big) {
        case 0: return {bv, (m[1][2] - m[2][1]) * mult, (m[2][0] - m[0][2]) * mult, (m[0][1] - m[1][0]) * mult};
        case 1: return {(m[1][2] - m[2][1]) * mult, bv, (m[0][1] + m[1][0]) * mult, (m[2][0] + m[0][2]) * mult};
        case 2: return {(m[2][0] - m[0][2]) * mult, (m[0][1] + m[1][0]) * mult, bv, (m[1][2] + m[2][1]) * mult};
        default: return {(m[0][1] - m[1][0]) * mult, (m[2][0] + m[0][2]) * mult, (m[1][2] + m[2][1]) * mult, bv};
    }
}

// glm::mat4_cast (gtc/quaternion.inl:47-79) into the upper 3x3 of an identity.
hm::M4 mat4_cast(Quat q) {
    hm::M4 R = hm::identity();
    const float qxx = q.x * q.x, qyy = q.y * q.y, qzz = q.z * q.z;
    const float qxz = q.x * q.z, qxy = q.x * q.y, qyz = q.y * q.z;
    const float qwx = q.w * q.x, qwy = q.w * q.y, qwz = q.w * q.z;
    R.at(0, 0) = 1.0f - 2.0f * (qyy + qzz);
    R.at(0, 1) = 2.0f * (qxy + qwz);
    R.at(0, 2) = 2.0f * (qxz - qwy);
    R.at(1, 0) = 2.0f * (qxy - qwz);
    R.at(1, 1) = 1.0f - 2.0f * (qxx + qzz);
    R.at(1, 2) = 2.0f * (qyz + qwx);
    R.at(2, 0) = 2.0f * (qxz + qwy);
    R.at(2, 1) = 2.0f * (qyz - qwx);
    R.at(2, 2) = 1.0f - 2.0f * (qxx + qyy);
    return R;
}

// The camera's rotation as the viewer stores it: quat_cast of the basis
// (basis[c][r] = file rotation[r][c]).
Quat state_rotation(const gps_camera_state* s) {
    float b[3][3];
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) b[c][r] = s->rotation[3 * r + c];
    return quat_cast(b);
}

// glm::slerp (ext/quaternion_common.inl:41-73), epsilon = FLT_EPSILON.
Quat slerp(Quat x, Quat y, float a) {
    Quat z = y;
    float c = (x.w * y.w + x.x * y.x) + (x.y * y.y + x.z * y.z);  // dot (detail/type_quat.inl:17-24)
    if (c < 0.0f) {
        z = {-y.w, -y.x, -y.y, -y.z};
        c = -c;
    }
    if (c > 1.0f - 1.1920929e-7f) {
        auto mix = [](float p, float q, float t) { return p * (1.0f - t) + q * t; };
        return {mix(x.w, z.w, a), mix(x.x, z.x, a), mix(x.y, z.y, a), mix(x.z, z.z, a)};
    }
    const float ang = std::acos(c);
    const float s0 = std::sin((1.0f - a) * ang), s1 = std::sin(a * ang), sa = std::sin(ang);
    return {(s0 * x.w + s1 * z.w) / sa, (s0 * x.x + s1 * z.x) / sa, (s0 * x.y + s1 * z.y) / sa,
            (s0 * x.z + s1 * z.z) / sa};
}

void fill_frame(const float pos[3], Quat q, float fovy, int32_t width, int32_t height, int32_t ss, float zn, float zf,
                gps_frame_params* f) {
    std::memset(f, 0, sizeof(*f));
    hm::M4 model = hm::identity();
    model.at(3, 0) = pos[0];  // glm::translate(I, p): column 3 = p (exact)
    model.at(3, 1) = pos[1];
    model.at(3, 2) = pos[2];
    model = hm::mul(model, mat4_cast(q));
    hm::M4 flip = hm::identity();
    flip.at(1, 1) = -1.0f;
    flip.at(2, 2) = -1.0f;
    const hm::M4 view = hm::mul(flip, hm::affine_inverse(model));
    std::memcpy(f->view, view.m, 64);
    const float aspect = float(width) / height;
    const hm::M4 proj = hm::perspective(fovy, aspect, zn, zf);
    std::memcpy(f->proj, proj.m, 64);
    hm::frustum_planes(fovy, view, aspect, zn, zf, f->frustum_planes);
    f->fovy_degrees = fovy;
    f->near_plane = zn;
    f->far_plane = zf;
    f->width = width;
    f->height = height;
    f->supersampling_factor = ss;
    f->frame = 1;
    f->enable_occlusion_culling = 1;
    f->enable_hierarchical_culling = 1;
    f->use_unbiased_2d_splatting = 1;
    f->reduce_point_count = 1;
}

}  // namespace

extern "C" {

gps_status gps_camera_path_load(const char* path, gps_camera_state* out, int32_t capacity, int32_t* count) {
    if (!path || !count || capacity < 0 || (capacity > 0 && !out)) return GPS_ERR_INVALID_ARG;
    *count = 0;
    std::ifstream in(path, std::ios::binary);
    if (!in) return GPS_ERR_IO;
    std::stringstream ss;
    ss << in.rdbuf();
    const std::string text = ss.str();
    JParser jp{text.data(), text.data() + text.size()};
    const JVal root = jp.value();
    jp.ws();
    if (!jp.ok || jp.p != jp.end || root.kind != JVal::ARR) return GPS_ERR_IO;
    std::vector<gps_camera_state> states(root.arr.size());
    for (size_t i = 0; i < root.arr.size(); ++i)
        if (!to_state(root.arr[i], &states[i])) return GPS_ERR_IO;
    // loadCameras sorts by id (camera_loader.cu:140-145); equal ids keep file order.
    std::stable_sort(states.begin(), states.end(),
                     [](const gps_camera_state& a, const gps_camera_state& b) { return a.id < b.id; });
    *count = (int32_t)states.size();
    const size_t n = std::min<size_t>(states.size(), (size_t)capacity);
    if (n) std::memcpy(out, states.data(), n * sizeof(gps_camera_state));
    return GPS_OK;
}

gps_status gps_camera_path_save(const char* path, const gps_camera_state* states, int32_t count) {
    if (!path || count < 0 || (count > 0 && !states)) return GPS_ERR_INVALID_ARG;
    std::FILE* f = std::fopen(path, "wb");
    if (!f) return GPS_ERR_IO;
    // nlohmann's dump(4) layout with its (alphabetical) key order; floats
    // with 9 significant digits, which read back to the same float.
    std::fprintf(f, "[");
    for (int32_t i = 0; i < count; ++i) {
        const gps_camera_state& s = states[i];
        std::string name;
        for (const char* c = s.img_name; *c; ++c) {
            if (*c == '"' || *c == '\\') name += '\\';
            name += *c;
        }
        std::fprintf(f, "%s\n    {\n        \"fx\": %.9g,\n        \"fy\": %.9g,\n        \"height\": %d,\n"
                        "        \"id\": %d,\n        \"img_name\": \"%s\",\n        \"position\": [\n"
                        "            %.9g,\n            %.9g,\n            %.9g\n        ],\n"
                        "        \"rotation\": [\n",
                     i ? "," : "", s.fx, s.fy, s.height, s.id, name.c_str(), s.position[0], s.position[1],
                     s.position[2]);
        for (int r = 0; r < 3; ++r)
            std::fprintf(f, "            [\n                %.9g,\n                %.9g,\n                %.9g\n"
                            "            ]%s\n",
                         s.rotation[3 * r], s.rotation[3 * r + 1], s.rotation[3 * r + 2], r < 2 ? "," : "");
        std::fprintf(f, "        ],\n        \"width\": %d\n    }", s.width);
    }
    std::fprintf(f, "%s]\n", count ? "\n" : "");
    return std::fclose(f) == 0 ? GPS_OK : GPS_ERR_IO;
}

float gps_camera_fovy(const gps_camera_state* s) {
    if (!s || !(s->fy > 0.0f) || s->height <= 0) return 0.0f;
    const float fovr = (float)(2.0 * (double)std::atan(0.5f * (float)s->height / s->fy));
    return fovr * 57.2957795f;
}

gps_status gps_camera_resolution(const gps_camera_state* s, int32_t width, int32_t* out_width, int32_t* out_height) {
    if (!s || !out_width || !out_height || width <= 0 || s->width <= 0 || s->height <= 0) return GPS_ERR_INVALID_ARG;
    const float aspect_yx = float(s->height) / s->width;
    *out_width = (int32_t)std::round((float)width);
    *out_height = (int32_t)std::round((float)width * aspect_yx);
    return *out_height > 0 ? GPS_OK : GPS_ERR_INVALID_ARG;
}

gps_status gps_frame_params_from_camera(const gps_camera_state* s, int32_t width, int32_t height,
                                        float fovy_degrees, int32_t supersampling_factor, float near_plane,
                                        float far_plane, gps_frame_params* out) {
    if (!s || !out || width <= 0 || height <= 0 || supersampling_factor < 1 || !(far_plane > near_plane))
        return GPS_ERR_INVALID_ARG;
    const float fovy = fovy_degrees > 0.0f ? fovy_degrees : gps_camera_fovy(s);
    if (!(fovy > 0.0f)) return GPS_ERR_INVALID_ARG;
    fill_frame(s->position, state_rotation(s), fovy, width, height, supersampling_factor, near_plane, far_plane, out);
    return GPS_OK;
}

gps_status gps_camera_path_frame(const gps_camera_state* states, int32_t count, float time_s, float fovy_degrees,
                                 int32_t width, int32_t height, int32_t supersampling_factor, float near_plane,
                                 float far_plane, gps_frame_params* out, int32_t* finished) {
    if (!states || count < 1 || !out || width <= 0 || height <= 0 || supersampling_factor < 1 ||
        !(far_plane > near_plane) || !(time_s >= 0.0f))
        return GPS_ERR_INVALID_ARG;
    const float interval = 1.0f / 24.0f;
    const float duration = interval * std::max((float)(count - 1), 0.0f);
    const float elapsed = std::min(std::max(time_s, 0.0f), duration);
    // keyframe pair (the reference would read past the last keyframe when
    // elapsed reaches the end exactly; the last pair is used instead)
    int32_t i = (int32_t)std::floor(elapsed / interval);
    i = std::max(0, std::min(i, count - 2));
    const gps_camera_state& a = states[count > 1 ? i : 0];
    const gps_camera_state& b = states[count > 1 ? i + 1 : 0];
    const float t = std::min(std::max(elapsed / interval - (float)i, 0.0f), 1.0f);
    float pos[3];
    for (int k = 0; k < 3; ++k) pos[k] = a.position[k] * (1.0f - t) + b.position[k] * t;
    const Quat q = slerp(state_rotation(&a), state_rotation(&b), t);
    const float fovy = fovy_degrees > 0.0f ? fovy_degrees : gps_camera_fovy(&states[0]);
    if (!(fovy > 0.0f)) return GPS_ERR_INVALID_ARG;
    fill_frame(pos, q, fovy, width, height, supersampling_factor, near_plane, far_plane, out);
    if (finished) *finished = elapsed >= duration ? 1 : 0;
    return GPS_OK;
}

}  // extern "C"
