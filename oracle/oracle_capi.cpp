// oracle_capi.cpp — C entry points of the CPU oracle (TEST INFRASTRUCTURE ONLY).
//
// Loaded by tests/ (ctypes), __graft_entry__.smoke() and bench.py's cpu_baseline leg, never by
// the product.  One handle wraps an oracle::World<float> or oracle::World<double>.
#include <cstring>
#include <exception>
#include <string>

#include "alien_oracle.hpp"

namespace {
thread_local std::string g_err;

struct Handle {
    int precision;  // 32 or 64
    oracle::World<float>* f = nullptr;
    oracle::World<double>* d = nullptr;
};

template <class Fn> int guard(Fn&& fn) {
    try {
        fn();
        return 0;
    } catch (std::exception const& e) {
        g_err = e.what();
        return -1;
    }
}
template <class Fn> auto dispatch(Handle* h, Fn&& fn) {
    if (h->precision == 32) return fn(*h->f);
    return fn(*h->d);
}
}  // namespace

extern "C" {

const char* oracle_last_error() { return g_err.c_str(); }

void* oracle_create(int precision, int W, int H, const alien_params* params) {
    Handle* h = new Handle();
    h->precision = precision;
    int rc = guard([&] {
        if (precision == 32)
            h->f = new oracle::World<float>(W, H, *params);
        else if (precision == 64)
            h->d = new oracle::World<double>(W, H, *params);
        else
            throw std::invalid_argument("precision must be 32 or 64");
    });
    if (rc) {
        delete h;
        return nullptr;
    }
    return h;
}

void oracle_destroy(void* hv) {
    Handle* h = static_cast<Handle*>(hv);
    delete h->f;
    delete h->d;
    delete h;
}

void oracle_set_threads(void* hv, int n) {
    dispatch(static_cast<Handle*>(hv), [&](auto& w) {
        w.numThreads = n < 1 ? 1 : n;
        return 0;
    });
}

void oracle_set_seed(void* hv, uint64_t seed) {
    dispatch(static_cast<Handle*>(hv), [&](auto& w) {
        w.seed = seed;
        return 0;
    });
}

// one uniform draw of the counter-based generator (for its known-answer test)
float oracle_uniform_draw(uint64_t seed, uint64_t timestep, uint64_t id, uint32_t purpose) {
    return oracle::uniformDraw(seed, timestep, id, purpose);
}

// Philox4x32-10 on a raw counter and key (for the generator's published known-answer vectors)
void oracle_philox4x32_10(uint32_t* counter, uint32_t k0, uint32_t k1) { oracle::philox4x32_10(counter, k0, k1); }

int oracle_set_params(void* hv, const alien_params* params) {
    return guard([&] { dispatch(static_cast<Handle*>(hv), [&](auto& w) { w.setParams(*params); return 0; }); });
}

int oracle_set_data(void* hv, const alien_cell_to* c, uint64_t n, const alien_particle_to* p, uint64_t np) {
    return guard([&] { dispatch(static_cast<Handle*>(hv), [&](auto& w) { w.setData(c, n, p, np); return 0; }); });
}

int oracle_get_counts(void* hv, uint64_t* n, uint64_t* np) {
    return dispatch(static_cast<Handle*>(hv), [&](auto& w) {
        *n = w.cells.size();
        *np = w.parts.size();
        return 0;
    });
}

int oracle_get_data(void* hv, alien_cell_to* c, uint64_t cap, uint64_t* n, alien_particle_to* p, uint64_t capP,
                    uint64_t* np) {
    return guard([&] {
        dispatch(static_cast<Handle*>(hv), [&](auto& w) {
            std::vector<alien_cell_to> oc;
            std::vector<alien_particle_to> op;
            w.getData(oc, op);
            if (oc.size() > cap || op.size() > capP) throw std::length_error("buffer too small");
            std::memcpy(c, oc.data(), oc.size() * sizeof(alien_cell_to));
            std::memcpy(p, op.data(), op.size() * sizeof(alien_particle_to));
            *n = oc.size();
            *np = op.size();
            return 0;
        });
    });
}

int oracle_step(void* hv, uint64_t steps) {
    return guard([&] {
        dispatch(static_cast<Handle*>(hv), [&](auto& w) {
            for (uint64_t s = 0; s < steps; ++s) w.step();
            return 0;
        });
    });
}

uint64_t oracle_get_timestep(void* hv) {
    return dispatch(static_cast<Handle*>(hv), [&](auto& w) { return w.timestep; });
}
void oracle_set_timestep(void* hv, uint64_t t) {
    dispatch(static_cast<Handle*>(hv), [&](auto& w) {
        w.timestep = t;
        return 0;
    });
}

int oracle_get_totals(void* hv, alien_totals* out) {
    return dispatch(static_cast<Handle*>(hv), [&](auto& w) {
        *out = w.totals();
        return 0;
    });
}

// cell map of the current state: per cell in map order, its bin and id
int oracle_get_map(void* hv, int32_t* bins, uint64_t* ids, uint64_t cap, uint64_t* n, int32_t* binSize,
                   int32_t* nbx, int32_t* nby) {
    return guard([&] {
        dispatch(static_cast<Handle*>(hv), [&](auto& w) {
            w.computeGeometry();
            w.sortCells();
            if (w.cells.size() > cap) throw std::length_error("buffer too small");
            for (size_t i = 0; i < w.cells.size(); ++i) {
                bins[i] = w.cellBinOfIndex[i];
                ids[i] = w.cells[i].id;
            }
            *n = w.cells.size();
            *binSize = w.binSize;
            *nbx = w.nbx;
            *nby = w.nby;
            return 0;
        });
    });
}

int oracle_neighbors(void* hv, uint64_t* offsets, uint64_t offCap, uint64_t* ids, uint64_t idsCap, uint64_t* numIds) {
    return guard([&] {
        dispatch(static_cast<Handle*>(hv), [&](auto& w) {
            std::vector<uint64_t> off, nid;
            w.neighborLists(off, nid);
            if (off.size() > offCap || nid.size() > idsCap) throw std::length_error("buffer too small");
            std::memcpy(offsets, off.data(), off.size() * sizeof(uint64_t));
            std::memcpy(ids, nid.data(), nid.size() * sizeof(uint64_t));
            *numIds = nid.size();
            return 0;
        });
    });
}

// scalar helpers, exposed so the tests can check them against closed forms
void oracle_sph_kernel(int precision, double q, double* w, double* dw) {
    if (precision == 32) {
        *w = oracle::calcKernel<float>(float(q));
        *dw = oracle::calcKernel_d<float>(float(q));
    } else {
        *w = oracle::calcKernel<double>(q);
        *dw = oracle::calcKernel_d<double>(q);
    }
}

double oracle_angle_of_vector(int precision, double x, double y) {
    if (precision == 32) return oracle::angleOfVector<float>(oracle::V2<float>{float(x), float(y)});
    return oracle::angleOfVector<double>(oracle::V2<double>{x, y});
}

double oracle_subtract_angle(int precision, double a, double b) {
    if (precision == 32) return oracle::subtractAngle<float>(float(a), float(b));
    return oracle::subtractAngle<double>(a, b);
}

}  // extern "C"

extern "C" {
// Math::crossing of segment (s, e) with (os, oe), exposed for its closed-form test
int oracle_crossing(int precision, const double* s, const double* e, const double* os, const double* oe) {
    if (precision == 32) {
        using V = oracle::V2<float>;
        return oracle::crossing<float>(V{float(s[0]), float(s[1])}, V{float(e[0]), float(e[1])},
                                       V{float(os[0]), float(os[1])}, V{float(oe[0]), float(oe[1])});
    }
    using V = oracle::V2<double>;
    return oracle::crossing<double>(V{s[0], s[1]}, V{e[0], e[1]}, V{os[0], os[1]}, V{oe[0], oe[1]});
}
}  // extern "C"
