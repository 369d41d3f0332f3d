/*
 * MI355XSimulator.hh : C++ host class over the C-ABI with the method set of the
 * reference's simulator protocol (sysrap/SSimulator.h), so it can be selected
 * where G4CXOpticks::CreateSimulator (g4cx/G4CXOpticks.cc:416-442) picks
 * CSGOptiX / SOpticksClientSimulator / SOpticksDummySimulator.
 *
 *   SSimulator method            here
 *   simulate_launch / launch     propagate the collected gensteps of the current event
 *   simulate(eventID, reset)     same, returns device ms; reset=true releases the event
 *   simtrace(eventID) / simtrace_launch
 *                                the collected OpticksGenstep_FRAME gensteps traced
 *                                as a simtrace grid (QSim::simtrace,
 *                                qudarap/QSim.cc:876-910), returns device ms;
 *                                records (n,4,4) via simtrace_records()
 *   render / render_launch       not part of this path: return -1 (unsupported)
 *   reset(eventID)               drop gensteps and the event's outputs
 *   desc()                       one-line description
 *
 * Gensteps are collected with add_genstep (the role of SEvt::addGenstep,
 * sysrap/SEvt.cc:1564); after simulate the hits/photons/seq stay readable until
 * reset, like the SEvt fold kept by simulate(eventID, reset=false).
 * The class is header-only and depends only on include/opticks_amd.h
 * (and mi_npy.hh for Load).
 *
 * MI355XSimulator::Load(base) reads the reference's persisted geometry and
 * physics (CSGFoundry::Load, CSG/CSGFoundry.h:125-133: <base>/CSGFoundry/
 * {solid,prim,node,plan,tran,itra,inst}.npy and SSim/stree/standard/
 * {bnd,optical,icdf}.npy with bnd_meta.txt) and creates the simulator.
 */
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/opticks_amd.h"
#include "mi_npy.hh"

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

    /* geometry + physics from a reference-layout folder (base or base/CSGFoundry) */
    static MI355XSimulator* Load(const std::string& base, const mi_config* cfg = nullptr)
    {
        std::string cf = base + "/CSGFoundry";
        if (!mi_npy::exists(cf + "/solid.npy")) cf = base;
        const std::string sd = cf + "/SSim/stree/standard";
        Arrays a;
        a.solid = mi_npy::load(cf + "/solid.npy").as_i32();
        a.prim = mi_npy::load(cf + "/prim.npy").as_f32();
        a.node = mi_npy::load(cf + "/node.npy").as_f32();
        a.inst = mi_npy::load(cf + "/inst.npy").as_f32();
        if (mi_npy::exists(cf + "/itra.npy")) a.itra = mi_npy::load(cf + "/itra.npy").as_f32();
        if (mi_npy::exists(cf + "/plan.npy")) a.plan = mi_npy::load(cf + "/plan.npy").as_f32();
        mi_npy::Array bnd = mi_npy::load(sd + "/bnd.npy");
        if (bnd.shape.size() != 5) throw std::runtime_error("bnd.npy: expected (nbnd,4,2,nwl,4)");
        a.bnd = bnd.as_f32();
        a.optical = mi_npy::load(sd + "/optical.npy").as_i32();
        int32_t n_icdf = 0, n_layer = 0;
        if (mi_npy::exists(sd + "/icdf.npy")) {
            mi_npy::Array ic = mi_npy::load(sd + "/icdf.npy");
            a.icdf = ic.as_f32();
            n_layer = (int32_t)ic.shape[0];
            n_icdf = (int32_t)(ic.size() / n_layer);
        }
        mi_geometry g = mi_geometry();
        g.node = a.node.data(); g.n_node = (int64_t)a.node.size() / 16;
        g.prim = a.prim.data(); g.n_prim = (int64_t)a.prim.size() / 16;
        g.solid = a.solid.data(); g.n_solid = (int64_t)a.solid.size() / 12;
        g.itra = a.itra.empty() ? nullptr : a.itra.data(); g.n_tran = (int64_t)a.itra.size() / 16;
        g.plan = a.plan.empty() ? nullptr : a.plan.data(); g.n_plan = (int64_t)a.plan.size() / 4;
        g.inst = a.inst.data(); g.n_inst = (int64_t)a.inst.size() / 16;
        mi_physics p = mi_physics();
        p.bnd = a.bnd.data(); p.n_bnd = (int32_t)bnd.shape[0]; p.n_wl = (int32_t)bnd.shape[3];
        p.wl_low = (float)mi_npy::meta_value(sd + "/bnd_meta.txt", "domain_low", 60.0);
        p.wl_step = (float)mi_npy::meta_value(sd + "/bnd_meta.txt", "domain_step", 1.0);
        p.optical = a.optical.data();
        p.icdf = a.icdf.empty() ? nullptr : a.icdf.data();
        p.n_icdf_layer = n_layer; p.n_icdf = n_icdf;
        p.icdf_hd_factor = n_layer == 3 ? 20 : 0;          /* qudarap/QSim.cc:184 */
        return new MI355XSimulator(g, p, cfg);             /* arrays are copied to the device at create */
    }

    /* genstep collection: one quad6 (6x4 floats/ints) per call, or many */
    void add_genstep(const float* quad6) { gensteps_.insert(gensteps_.end(), quad6, quad6 + 24); }
    void add_gensteps(const float* q, int64_t n) { gensteps_.insert(gensteps_.end(), q, q + 24 * n); }
    int64_t num_genstep() const { return (int64_t)gensteps_.size() / 24; }

    /* SSimulator protocol */
    virtual double render_launch() { return -1.0; }
    virtual double simtrace_launch() { return run_simtrace(); }
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
    virtual double simtrace(int eventID)
    {
        event_id_ = eventID;
        return run_simtrace();
    }
    virtual double render(const char* = nullptr) { return -1.0; }
    virtual void reset(int eventID)
    {
        mi_sim_reset(sim_, eventID);
        gensteps_.clear();
        out_ = mi_event_out();
        simtrace_ = nullptr;
        n_simtrace_ = 0;
    }

    /* event outputs (valid until reset) */
    int64_t num_photon() const { return out_.n_photon; }
    int64_t num_hit() const { return out_.n_hit; }
    const float* photon() const { return out_.photon; }      /* (n,4,4) sphoton */
    const uint64_t* seq() const { return out_.seq; }         /* (n,2,2) seqhis, seqbnd */
    const float* hit() const { return out_.hit; }            /* (m,4,4) */
    int64_t num_simtrace() const { return n_simtrace_; }
    const float* simtrace_records() const { return simtrace_; }   /* (n,4,4) sevent simtrace layout */
    mi_sim* handle() { return sim_; }

private:
    struct Arrays {
        std::vector<int32_t> solid, optical;
        std::vector<float> prim, node, inst, itra, plan, bnd, icdf;
    };

    double run(int eventID)
    {
        out_ = mi_event_out();
        int rc = mi_sim_simulate(sim_, eventID, gensteps_.empty() ? nullptr : gensteps_.data(), num_genstep(), &out_);
        if (rc == MI_NOTHING_TO_DO) return 0.0;
        if (rc != MI_OK) throw std::runtime_error(std::string("mi_sim_simulate: ") + mi_last_error());
        return out_.device_ms;
    }

    double run_simtrace()
    {
        simtrace_ = nullptr;
        n_simtrace_ = 0;
        double ms = 0.0;
        int rc = mi_sim_simtrace(sim_, gensteps_.empty() ? nullptr : gensteps_.data(), num_genstep(), &simtrace_,
                                 &n_simtrace_, &ms);
        if (rc == MI_NOTHING_TO_DO) return -1.0;                /* the reference's dt = -1 without gensteps */
        if (rc != MI_OK) throw std::runtime_error(std::string("mi_sim_simtrace: ") + mi_last_error());
        return ms;
    }

    mi_sim* sim_ = nullptr;
    const float* simtrace_ = nullptr;
    int64_t n_simtrace_ = 0;
    std::vector<float> gensteps_;
    mi_event_out out_ = mi_event_out();
    int event_id_ = 0;
    std::string desc_;
};
