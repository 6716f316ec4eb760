// The C ABI's handle types, shared by the ABI and the solve server.
#pragma once

#include <memory>

#include "sf_device.hpp"

using namespace sf;

struct sf_context {
    std::shared_ptr<Context> ctx;  // shared with every object that launches on its stream
};
struct sf_tet_mesh {
    TetMesh mesh;
};
struct sf_quad_mesh {
    QuadMesh mesh;
};
struct sf_model {
    Model model;
    sf_quad_mesh view;  // the mesh as a handle
};
struct sf_csr {
    std::shared_ptr<Context> ctx;
    DeviceSystem system;
};
struct sf_multigrid {
    std::shared_ptr<Context> ctx;
    std::unique_ptr<DeviceMultigrid> mg;
    DeviceBuffer<double> in, out;  // apply()'s device copies, kept between calls
};
struct sf_multigrid_setup {
    std::shared_ptr<Context> ctx;
    std::unique_ptr<MultigridSetup> setup;
};
struct sf_pcg {
    std::unique_ptr<DevicePcg> pcg;
};

struct sf_pcg_batch {
    std::unique_ptr<DevicePcgBatch> batch;
};

struct sf_static_study {
    std::shared_ptr<Context> ctx;
    Model model;
    FreeProblem problem;
    PrescribedRhs prescribed_rhs;  // what eliminating the prescribed components does to a load
    sf_csr system;
    sf_multigrid mg;
    std::unique_ptr<sf_pcg> pcg;
    sf_report report{};
};

struct sf_error_estimation {
    std::shared_ptr<Context> ctx;
    std::unique_ptr<DeviceErrorEstimation> estimation;
};
struct sf_recovery {
    std::shared_ptr<Context> ctx;
    std::unique_ptr<DeviceRecovery> recovery;
};

struct sf_study_builder {
    std::shared_ptr<Context> ctx;
    Model model;
    FreeProblem problem;
    AssemblyPattern pattern;
    DeviceAssembly assembly;
    std::unique_ptr<sf_static_study> study;  // the system being assembled
    sf_pcg_options pcg{};
    bool use_multigrid = false;
    SetupOptions mg_options;
    size_t next = 0;
    bool prescribed_applied = false;
    std::unique_ptr<MultigridSetup> setup;
    bool setup_finished = false;
    double assembly_seconds = 0, preconditioner_seconds = 0;
};

