// The device side of the multigrid setup: a level's power iteration, smoothed prolongator,
// A P and Galerkin product on the GPU, bit-identical to the host phases they replace.
#pragma once

#include <memory>

#include "sf_device.hpp"
#include "sf_setup.hpp"

namespace sf {

class DeviceSetup final : public SetupAccelerator {
   public:
    // Levels of at least `min_rows` block rows run their numeric phases here. `level0_values`,
    // when given, is the finest matrix's values already on the device in block-row order (the
    // study's assembled stiffness); the setup borrows it while level 0 is built.
    // With `aggregate_on_device`, the seed pass of aggregation runs here too (in rounds, with the
    // serial pass's result).
    DeviceSetup(std::shared_ptr<Context> ctx, size_t min_rows, const double *level0_values, bool aggregate_on_device = false);
    ~DeviceSetup() override;

    bool takes(int level, const HostBsr<double> &a) const override;
    void begin(int level, const HostBsr<double> &a, const std::vector<double> &inverse_diagonal, const double *device_values) override;
    double spectral(std::vector<double> &power, int iterations) override;
    bool smooth(const std::vector<u32> &aggregate, u32 count, const std::vector<int64_t> &ns, const std::vector<u32> &nc,
                const std::vector<double> &tentative, double omega, HostBsr<double> &p) override;
    bool product(u32 count) override;
    bool galerkin(u32 count, const std::vector<int64_t> &pcol_starts, const std::vector<std::pair<u32, int64_t>> &entries,
                  HostBsr<double> &coarse) override;
    std::shared_ptr<const ValueSource> prolongator_values() override;
    std::shared_ptr<const ValueSource> coarse_values() override;
    void product_values(HostBsr<double> &ap) override;
    void end_level() override;
    bool seed(int level, const std::vector<int64_t> &starts, const std::vector<u32> &columns, std::vector<u32> &aggregate,
              std::vector<char> &seeded, u32 &count) override;
    bool factor(int level, size_t n, double *a) override;

   private:
    struct Overflow {};
    void check_overflow(const char *phase);
    void clear_overflow();

    std::shared_ptr<Context> ctx_;
    size_t min_rows_;
    const double *level0_values_;
    bool aggregate_on_device_;
    int level_ = -1, rb_ = 3;
    size_t rows_ = 0;
    DeviceBuffer<int64_t> a_starts_, p_starts_, ap_starts_;
    DeviceBuffer<u32> a_cols_, p_cols_, ap_cols_, overflow_;
    DeviceBuffer<double> a_vals_, dinv_, ap_vals_, x_, y_;
    std::shared_ptr<DeviceBuffer<double>> p_vals_;  // shared with the level's prolongator source
    u32 slot_limit_ = 1024;                          // output blocks per row the kernels take
    size_t p_blocks_ = 0;
    std::vector<int64_t> ap_host_starts_;  // the pattern of the last A P
    std::shared_ptr<const ValueSource> coarse_vals_;  // the last Galerkin product's values
    const double *a_vals_ptr_ = nullptr;
    std::vector<double> host_y_;
};

// Block values on the device as a value source: borrowed (a study's assembled values, which the
// study keeps while the hierarchy holding this source lives) or shared (a prolongator built on
// the device).
// Output blocks per row above which a device setup row takes the host path (0: the kernels' limit).
std::atomic<u32> &setup_slot_limit();

class DeviceValues final : public ValueSource {
   public:
    DeviceValues(std::shared_ptr<Context> ctx, const double *values, size_t blocks, size_t block_len);
    DeviceValues(std::shared_ptr<Context> ctx, std::shared_ptr<DeviceBuffer<double>> owned, size_t blocks, size_t block_len);
    void values(double *out) const override;
    void blocks(const std::vector<int64_t> &slots, double *out) const override;
    void block_squares(double *out) const override;
    const double *device() const override { return values_; }

   private:
    std::shared_ptr<Context> ctx_;
    std::shared_ptr<DeviceBuffer<double>> owned_;
    const double *values_;
    size_t blocks_, len_;
};

}  // namespace sf
