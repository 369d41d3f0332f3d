/*
 * MI355XSimulator.hh : C++ host class over the C-ABI with the method set of the
 * reference's simulator protocol (sysrap/SSimulator.h), so it can be selected
 * where G4CXOpticks::CreateSimulator (g4cx/G4CXOpticks.cc:416-442) picks
 * CSGOptiX / SOpticksClientSimulator / SOpticksDummySimulator.
 *
 *   SSimulator method            here
 *   simulate_launch / launch     propagate the collected gensteps of the current event
 *   simulate(eventID, reset)     same, returns device ms; reset=true releases the event
 *   simtrace / simtrace_launch   not part of this path: return -1 (unsupported)
 *   render / render_launch       not part of this path: return -1 (unsupported)
 *   reset(eventID)               drop gensteps and the event's outputs
 *   desc()                       one-line description
 *
 * Gensteps are collected with add_genstep (the role of SEvt::addGenstep,
 * sysrap/SEvt.cc:1564); after simulate the hits/photons/seq stay readable until
 * reset, like the SEvt fold kept by simulate(eventID, reset=false).
 * The class is header-only and depends only on include/opticks_amd.h.
 */
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/opticks_amd.h"

class MI355XSimulator {
public:
    MI355XSimulator(const mi_geometry& geo, const mi_physics& phys, const mi_config* cfg = nullptr)
    {
        int rc = mi_sim_create(&geo, &phys, cfg, &sim_);
        if (rc != MI_OK) throw std::runtime_error(std::string("mi_sim_create: ") + mi_last_error());
        desc_ = std::string("MI355XSimulator(") + mi_version() + ")";
    }
    virtual ~MI355XSimulator() { mi_sim_destroy(sim_); }
    MI355XSimulator(const MI355XSimulator&) = delete;
    MI355XSimulator& operator=(const MI355XSimulator&) = delete;

    /* genstep collection: one quad6 (6x4 floats/ints) per call, or many */
    void add_genstep(const float* quad6) { gensteps_.insert(gensteps_.end(), quad6, quad6 + 24); }
    void add_gensteps(const float* q, int64_t n) { gensteps_.insert(gensteps_.end(), q, q + 24 * n); }
    int64_t num_genstep() const { return (int64_t)gensteps_.size() / 24; }

    /* SSimulator protocol */
    virtual double render_launch() { return -1.0; }
    virtual double simtrace_launch() { return -1.0; }
    virtual double simulate_launch() { return run(event_id_); }
    virtual double launch() { return run(event_id_); }
    virtual const char* desc() const { return desc_.c_str(); }
    virtual double simulate(int eventID, bool reset_ = false)
    {
        event_id_ = eventID;
        double ms = run(eventID);
        if (reset_) reset(eventID);
        return ms;
    }
    virtual double simtrace(int) { return -1.0; }
    virtual double render(const char* = nullptr) { return -1.0; }
    virtual void reset(int eventID)
    {
        mi_sim_reset(sim_, eventID);
        gensteps_.clear();
        out_ = mi_event_out();
    }

    /* event outputs (valid until reset) */
    int64_t num_photon() const { return out_.n_photon; }
    int64_t num_hit() const { return out_.n_hit; }
    const float* photon() const { return out_.photon; }      /* (n,4,4) sphoton */
    const uint64_t* seq() const { return out_.seq; }         /* (n,2,2) seqhis, seqbnd */
    const float* hit() const { return out_.hit; }            /* (m,4,4) */
    mi_sim* handle() { return sim_; }

private:
    double run(int eventID)
    {
        out_ = mi_event_out();
        int rc = mi_sim_simulate(sim_, eventID, gensteps_.empty() ? nullptr : gensteps_.data(), num_genstep(), &out_);
        if (rc == MI_NOTHING_TO_DO) return 0.0;
        if (rc != MI_OK) throw std::runtime_error(std::string("mi_sim_simulate: ") + mi_last_error());
        return out_.device_ms;
    }

    mi_sim* sim_ = nullptr;
    std::vector<float> gensteps_;
    mi_event_out out_ = mi_event_out();
    int event_id_ = 0;
    std::string desc_;
};
