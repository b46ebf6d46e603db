// settings.cpp — ALIEN's .settings.json for the physics path (alien_settings_read / _write).
//
// ALIEN writes its settings with boost::property_tree (AuxiliaryDataParserService.cpp:1100-1116
// for the general part, :30-763 for the simulation parameters): nested JSON objects whose leaves
// are all strings, colour vectors as sibling keys "name[0]".."name[6]", and a parameter missing
// from the file takes its default (ParameterParser::encodeDecode).  This file reads the physics
// subset of such a file into alien_params / alien_world_desc, reports which processes of the file
// this path does not simulate, and writes the subset back in the same layout.
#include <cmath>
#include <limits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/alien_hip.h"
#include "capi_guard.hpp"
#include "worldset.hpp"

using alien::Error;

namespace {

// ---- a small JSON reader: objects, arrays, strings, numbers, literals -------------------------
// Leaves are collected as dotted paths ("a.b.c") with their text; array elements get "[i]".
class JsonReader {
public:
    JsonReader(const char* s, size_t n) : s_(s), n_(n) {}
    std::map<std::string, std::string> parse() {
        std::map<std::string, std::string> out;
        ws();
        value("", out, 0);
        ws();
        if (p_ != n_) fail("trailing characters");
        return out;
    }

private:
    const char* s_;
    size_t n_, p_ = 0;
    [[noreturn]] void fail(const char* what) {
        throw Error(ALIEN_E_ARG, std::string("settings JSON: ") + what + " at offset " + std::to_string(p_));
    }
    void ws() {
        while (p_ < n_ && (s_[p_] == ' ' || s_[p_] == '\n' || s_[p_] == '\r' || s_[p_] == '\t')) ++p_;
    }
    bool eat(char c) {
        ws();
        if (p_ < n_ && s_[p_] == c) {
            ++p_;
            return true;
        }
        return false;
    }
    std::string string() {
        if (!eat('"')) fail("expected a string");
        std::string r;
        while (true) {
            if (p_ >= n_) fail("unterminated string");
            char c = s_[p_++];
            if (c == '"') break;
            if (c != '\\') {
                r += c;
                continue;
            }
            if (p_ >= n_) fail("unterminated escape");
            c = s_[p_++];
            switch (c) {
                case '"': r += '"'; break;
                case '\\': r += '\\'; break;
                case '/': r += '/'; break;
                case 'b': r += '\b'; break;
                case 'f': r += '\f'; break;
                case 'n': r += '\n'; break;
                case 'r': r += '\r'; break;
                case 't': r += '\t'; break;
                case 'u': {
                    if (p_ + 4 > n_) fail("short \\u escape");
                    const unsigned cp = unsigned(std::strtoul(std::string(s_ + p_, 4).c_str(), nullptr, 16));
                    p_ += 4;
                    if (cp < 0x80) {
                        r += char(cp);
                    } else if (cp < 0x800) {
                        r += char(0xC0 | (cp >> 6));
                        r += char(0x80 | (cp & 0x3F));
                    } else {
                        r += char(0xE0 | (cp >> 12));
                        r += char(0x80 | ((cp >> 6) & 0x3F));
                        r += char(0x80 | (cp & 0x3F));
                    }
                    break;
                }
                default: fail("bad escape");
            }
        }
        return r;
    }
    void value(const std::string& path, std::map<std::string, std::string>& out, int depth) {
        if (depth > 64) fail("nesting too deep");
        ws();
        if (p_ >= n_) fail("unexpected end");
        const char c = s_[p_];
        if (c == '{') {
            ++p_;
            if (eat('}')) return;
            do {
                ws();
                const std::string key = string();
                if (!eat(':')) fail("expected ':'");
                value(path.empty() ? key : path + "." + key, out, depth + 1);
            } while (eat(','));
            if (!eat('}')) fail("expected '}'");
        } else if (c == '[') {
            ++p_;
            if (eat(']')) return;
            int i = 0;
            do value(path + "[" + std::to_string(i++) + "]", out, depth + 1);
            while (eat(','));
            if (!eat(']')) fail("expected ']'");
        } else if (c == '"') {
            out[path] = string();
        } else {
            const size_t b = p_;
            while (p_ < n_ && std::strchr(",}] \n\r\t", s_[p_]) == nullptr) ++p_;
            if (p_ == b) fail("expected a value");
            out[path] = std::string(s_ + b, p_ - b);
        }
    }
};

constexpr const char* kSim = "simulation parameters.";

struct Reader {
    const std::map<std::string, std::string>& kv;
    bool has(const std::string& k) const { return kv.count(k) != 0; }
    template <class T> void num(const std::string& key, T& v) const {
        auto it = kv.find(key);
        if (it == kv.end()) return;  // missing: the default stays (ParameterParser::encodeDecode)
        const char* t = it->second.c_str();
        char* end = nullptr;
        const double d = std::strtod(t, &end);
        if (end == t) throw Error(ALIEN_E_ARG, "settings JSON: " + key + " is not a number");
        v = T(d);
    }
    void colors(const std::string& key, float* v) const {
        for (int c = 0; c < ALIEN_MAX_COLORS; ++c) num(key + "[" + std::to_string(c) + "]", v[c]);
    }
    bool flag(const std::string& key) const {
        auto it = kv.find(key);
        return it != kv.end() && it->second == "true";
    }
    double value(const std::string& key, double dflt) const {
        double d = dflt;
        num(key, d);
        return d;
    }
};

// ---- writer: nested objects from dotted paths, leaves as strings --------------------------------
struct Node {
    std::map<std::string, Node> children;
    std::vector<std::string> order;
    std::string text;
    bool leaf = false;
    Node& child(const std::string& k) {
        if (!children.count(k)) order.push_back(k);
        return children[k];
    }
};

void put(Node& root, const std::string& path, const std::string& text) {
    Node* n = &root;
    size_t b = 0;
    while (true) {
        const size_t e = path.find('.', b);
        n = &n->child(path.substr(b, e == std::string::npos ? std::string::npos : e - b));
        if (e == std::string::npos) break;
        b = e + 1;
    }
    n->leaf = true;
    n->text = text;
}

void emit(const Node& n, std::string& out, int indent) {
    if (n.leaf) {
        out += '"' + n.text + '"';
        return;
    }
    out += "{\n";
    for (size_t i = 0; i < n.order.size(); ++i) {
        out += std::string(size_t(indent + 4), ' ') + '"' + n.order[i] + "\": ";
        emit(n.children.at(n.order[i]), out, indent + 4);
        out += i + 1 < n.order.size() ? ",\n" : "\n";
    }
    out += std::string(size_t(indent), ' ') + "}";
}

std::string fmt(double v) {  // property_tree's float text: fixed, 8 decimals
    char b[128];
    std::snprintf(b, sizeof(b), "%.8f", v);
    return b;
}

}  // namespace

extern "C" {

int alien_settings_read(const char* json, size_t length, alien_world_desc* desc, alien_params* params,
                        uint64_t* timestep, uint32_t* unsimulated) {
    return alien::guarded([&] {
        if (!json || !desc || !params) throw Error(ALIEN_E_ARG, "null argument");
        const auto kv = JsonReader(json, length).parse();
        const Reader r{kv};
        // general settings (AuxiliaryDataParserService.cpp:1106-1112); world size defaults 1000 x 500
        std::memset(desc, 0, sizeof(*desc));
        desc->worldSizeX = 1000;
        desc->worldSizeY = 500;
        r.num("general.world size.x", desc->worldSizeX);
        r.num("general.world size.y", desc->worldSizeY);
        if (timestep) {
            uint64_t t = 0;
            r.num("general.time step", t);
            *timestep = t;
        }
        // the physics subset of the simulation parameters (:54-131, :231-233), ALIEN's defaults
        // for what the file leaves out; innerFriction is not part of the file (default 0.3)
        alien_params_default(params);
        alien_params& p = *params;
        const std::string S = kSim;
        r.num(S + "time step size", p.timestepSize);
        r.num(S + "motion.type", p.motionType);
        if (p.motionType == ALIEN_MOTION_FLUID) {
            r.num(S + "fluid.smoothing length", p.smoothingLength);
            r.num(S + "fluid.pressure strength", p.pressureStrength);
            r.num(S + "fluid.viscosity strength", p.viscosityStrength);
        } else {
            r.num(S + "motion.collision.max distance", p.cellMaxCollisionDistance);
            r.num(S + "motion.collision.repulsion strength", p.cellRepulsionStrength);
        }
        r.num(S + "friction", p.friction);
        r.num(S + "cell.max velocity", p.cellMaxVelocity);
        r.colors(S + "cell.max binding distance", p.cellMaxBindingDistance);
        r.num(S + "cell.min distance", p.cellMinDistance);
        r.colors(S + "cell.max force", p.cellMaxForce);
        r.num(S + "cell.max force decay probability", p.cellMaxForceDecayProb);
        r.colors(S + "cell.min energy", p.cellMinEnergy);
        r.num(S + "cell.fusion velocity", p.cellFusionVelocity);
        r.num(S + "cell.max binding energy", p.cellMaxBindingEnergy);
        r.colors(S + "radiation.absorption", p.radiationAbsorption);
        // cell radiation (:203-268): radiationCellAgeStrength is "radiation.factor"
        r.colors(S + "radiation.factor", p.radiationCellAgeStrength);
        r.num(S + "radiation.probability", p.radiationProb);
        r.num(S + "radiation.velocity multiplier", p.radiationVelocityMultiplier);
        r.num(S + "radiation.velocity perturbation", p.radiationVelocityPerturbation);
        r.colors(S + "high radiation.min cell energy", p.highRadiationMinCellEnergy);
        r.colors(S + "high radiation.factor", p.highRadiationFactor);
        for (int c = 0; c < ALIEN_MAX_COLORS; ++c)
            r.num(S + "radiation.min cell age[" + std::to_string(c) + "]", p.radiationMinCellAge[c]);
        // particle splitting (:662); ALIEN writes its "unbounded" default as the largest float
        r.colors(S + "particle.split energy", p.particleSplitEnergy);
        for (float& e : p.particleSplitEnergy)
            if (e >= 3.0e38f) e = std::numeric_limits<float>::infinity();
        // processes in the file that the physics path does not simulate
        uint32_t u = 0;
        if (r.value(S + "spots.num spots", 0.0) > 0.0) u |= ALIEN_SETTINGS_ZONES;
        if (r.value(S + "particle sources.num sources", 0.0) > 0.0) u |= ALIEN_SETTINGS_RADIATION_SOURCES;
        for (int c = 0; c < ALIEN_MAX_COLORS; ++c) {
            const std::string k = "[" + std::to_string(c) + "]";
            // the absorption penalties (SimulationParameters.h:83-84, ZoneValues.h:15-16)
            for (const char* pen : {"radiation.absorption velocity penalty", "radiation.absorption low velocity penalty",
                                    "radiation.absorption low connection penalty",
                                    "radiation.absorption low genome complexity penalty"})
                if (r.value(S + pen + k, 0.0) != 0.0) u |= ALIEN_SETTINGS_RADIATION;
            if (r.value(S + "cell.death probability" + k, 0.001) > 0.0) u |= ALIEN_SETTINGS_CELL_DEATH;
        }
        if (r.value(S + "cell.death consequences", 2.0) != 0.0) u |= ALIEN_SETTINGS_CELL_DEATH;
        if (r.value(S + "rigidity", 0.0) != 0.0) u |= ALIEN_SETTINGS_RIGIDITY;
        if (r.flag(S + "particle.transformation allowed")) u |= ALIEN_SETTINGS_PARTICLE_TRANSFORMATION;
        if (r.flag(S + "features.additional absorption control")) u |= ALIEN_SETTINGS_ABSORPTION_CONTROL;
        if (unsimulated) *unsimulated = u;
        alien::validate_params(p);
    });
}

int alien_settings_write(const alien_world_desc* desc, const alien_params* params, uint64_t timestep, char* buf,
                         size_t capacity, size_t* length) {
    return alien::guarded([&] {
        if (!desc || !params) throw Error(ALIEN_E_ARG, "null argument");
        const alien_params& p = *params;
        Node root;
        put(root, "general.time step", std::to_string(timestep));
        put(root, "general.world size.x", std::to_string(desc->worldSizeX));
        put(root, "general.world size.y", std::to_string(desc->worldSizeY));
        const std::string S = kSim;
        auto colors = [&](const std::string& key, const float* v) {
            for (int c = 0; c < ALIEN_MAX_COLORS; ++c) put(root, key + "[" + std::to_string(c) + "]", fmt(v[c]));
        };
        put(root, S + "time step size", fmt(p.timestepSize));
        put(root, S + "motion.type", std::to_string(p.motionType));
        if (p.motionType == ALIEN_MOTION_FLUID) {
            put(root, S + "fluid.smoothing length", fmt(p.smoothingLength));
            put(root, S + "fluid.pressure strength", fmt(p.pressureStrength));
            put(root, S + "fluid.viscosity strength", fmt(p.viscosityStrength));
        } else {
            put(root, S + "motion.collision.max distance", fmt(p.cellMaxCollisionDistance));
            put(root, S + "motion.collision.repulsion strength", fmt(p.cellRepulsionStrength));
        }
        put(root, S + "friction", fmt(p.friction));
        put(root, S + "cell.max velocity", fmt(p.cellMaxVelocity));
        colors(S + "cell.max binding distance", p.cellMaxBindingDistance);
        put(root, S + "cell.min distance", fmt(p.cellMinDistance));
        colors(S + "cell.max force", p.cellMaxForce);
        put(root, S + "cell.max force decay probability", fmt(p.cellMaxForceDecayProb));
        colors(S + "cell.min energy", p.cellMinEnergy);
        put(root, S + "cell.fusion velocity", fmt(p.cellFusionVelocity));
        // property_tree writes the largest float for "unbounded" (an infinite value has no JSON text)
        put(root, S + "cell.max binding energy", fmt(std::isinf(p.cellMaxBindingEnergy) ? 3.40282346638528859811704183484516925440e38
                                                                                       : double(p.cellMaxBindingEnergy)));
        colors(S + "radiation.absorption", p.radiationAbsorption);
        colors(S + "radiation.factor", p.radiationCellAgeStrength);
        put(root, S + "radiation.probability", fmt(p.radiationProb));
        put(root, S + "radiation.velocity multiplier", fmt(p.radiationVelocityMultiplier));
        put(root, S + "radiation.velocity perturbation", fmt(p.radiationVelocityPerturbation));
        colors(S + "high radiation.min cell energy", p.highRadiationMinCellEnergy);
        colors(S + "high radiation.factor", p.highRadiationFactor);
        for (int c = 0; c < ALIEN_MAX_COLORS; ++c) {
            put(root, S + "radiation.min cell age[" + std::to_string(c) + "]", std::to_string(p.radiationMinCellAge[c]));
            const float e = p.particleSplitEnergy[c];
            put(root, S + "particle.split energy[" + std::to_string(c) + "]",
                fmt(std::isinf(e) && e > 0 ? 3.40282346638528859811704183484516925440e38 : double(e)));
        }
        // what this path keeps fixed, so that ALIEN reads the file back to the same physics
        put(root, S + "cell.death consequences", "0");
        put(root, S + "spots.num spots", "0");
        put(root, S + "particle sources.num sources", "0");
        put(root, S + "rigidity", fmt(0.0));
        for (int c = 0; c < ALIEN_MAX_COLORS; ++c) {
            const std::string k = "[" + std::to_string(c) + "]";
            put(root, S + "cell.death probability" + k, fmt(0.0));
        }
        std::string out;
        emit(root, out, 0);
        out += "\n";
        if (length) *length = out.size();
        if (!buf || capacity < out.size() + 1)
            throw Error(ALIEN_E_CAPACITY, "settings buffer too small (the length was written)");
        std::memcpy(buf, out.c_str(), out.size() + 1);
    });
}

}  // extern "C"
