/*
 * SMI355XSimulator.hh : the MI355X simulate path as an implementation of the
 * reference's simulator protocol, struct SSimulator (sysrap/SSimulator.h:16-34),
 * beside its implementations CSGOptiX (CSGOptiX/CSGOptiX.cc:755,1218),
 * SOpticksClientSimulator and SOpticksDummySimulator. Selected where
 * G4CXOpticks::CreateSimulator (g4cx/G4CXOpticks.cc:416-442) picks one, it takes
 * CSGOptiX's place under G4CXOpticks::simulate (g4cx/G4CXOpticks.cc:470-489).
 *
 * Include path: the reference's sysrap (for SSimulator.h) and this directory.
 * Every method of the protocol is overridden (the compiler checks names and
 * signatures against SSimulator.h) and forwards to MI355XSimulator, the
 * header-only class over the C-ABI of include/opticks_amd.h:
 *
 *   render_launch / render(stem)   -1 (pixels are outside this path)
 *   simtrace_launch / simtrace(e)  OpticksGenstep_FRAME gensteps as a simtrace grid
 *   simulate_launch / launch       propagate the current event's gensteps
 *   simulate(eventID, reset)       same for eventID; device ms; reset=true drops the event
 *   reset(eventID)                 drop gensteps and outputs
 *   desc()                         one line
 *
 * Gensteps are added through sim() (MI355XSimulator::add_genstep(s)), the role of
 * SEvt::addGenstep (sysrap/SEvt.cc:1564); hits, photons and seq stay readable
 * through sim() until reset.
 */
#pragma once

#include "SSimulator.h"
#include "MI355XSimulator.hh"

struct SMI355XSimulator final : public SSimulator
{
    SMI355XSimulator(const mi_geometry& geo, const mi_physics& phys, const mi_config* cfg = nullptr)
        : owned_(new MI355XSimulator(geo, phys, cfg)) {}
    explicit SMI355XSimulator(MI355XSimulator* sim) : owned_(sim) {}
    ~SMI355XSimulator() override { delete owned_; }
    SMI355XSimulator(const SMI355XSimulator&) = delete;
    SMI355XSimulator& operator=(const SMI355XSimulator&) = delete;

    /* geometry + physics from a reference-layout CSGFoundry folder (CSGFoundry::Load) */
    static SMI355XSimulator* Load(const std::string& base, const mi_config* cfg = nullptr)
    {
        return new SMI355XSimulator(MI355XSimulator::Load(base, cfg));
    }

    MI355XSimulator& sim() { return *owned_; }
    const MI355XSimulator& sim() const { return *owned_; }

    /* SSimulator */
    double render_launch() override { return owned_->render_launch(); }
    double simtrace_launch() override { return owned_->simtrace_launch(); }
    double simulate_launch() override { return owned_->simulate_launch(); }
    double launch() override { return owned_->launch(); }
    const char* desc() const override { return owned_->desc(); }
    double simulate(int eventID, bool reset = false) override { return owned_->simulate(eventID, reset); }
    double simtrace(int eventID) override { return owned_->simtrace(eventID); }
    double render(const char* stem = nullptr) override { return owned_->render(stem); }
    void reset(int eventID) override { owned_->reset(eventID); }

private:
    MI355XSimulator* owned_;
};
