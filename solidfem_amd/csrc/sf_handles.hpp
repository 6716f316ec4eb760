// The C ABI's handle types, shared by the ABI and the solve server.
#pragma once

#include <memory>

#include "sf_device.hpp"

using namespace sf;

struct sf_context {
    std::unique_ptr<Context> ctx;
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
    sf_context *ctx = nullptr;
    DeviceSystem system;
};
struct sf_multigrid {
    sf_context *ctx = nullptr;
    std::unique_ptr<DeviceMultigrid> mg;
};
struct sf_multigrid_setup {
    sf_context *ctx = nullptr;
    std::unique_ptr<MultigridSetup> setup;
};
struct sf_pcg {
    std::unique_ptr<DevicePcg> pcg;
};

struct sf_pcg_batch {
    std::unique_ptr<DevicePcgBatch> batch;
};

struct sf_static_study {
    sf_context *ctx = nullptr;
    Model model;
    FreeProblem problem;
    sf_csr system;
    sf_multigrid mg;
    std::unique_ptr<sf_pcg> pcg;
    sf_report report{};
};

struct sf_study_builder {
    sf_context *ctx = nullptr;
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

