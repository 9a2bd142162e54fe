// GPU selection and memory budgeting: same interface as the reference
// (include/aa_device_info.hpp:61-130), backed by the HIP runtime.
// free_memory() reports what the pipeline may still allocate and
// request_memory() books an allocation against it; the strategy uses both.
#ifndef ASTRO_ACCELERATE_AA_DEVICE_INFO_HPP
#define ASTRO_ACCELERATE_AA_DEVICE_INFO_HPP

#include <cstddef>
#include <string>

namespace astroaccelerate {

class aa_device_info {
public:
    struct aa_card_info {
        int compute_capability_major = 0;
        int compute_capability_minor = 0;
        float clock_rate = 0;         // kHz
        float memory_clock_rate = 0;  // kHz
        float memory_bus_width = 0;   // bits
        int card_number = 0;
        int driver_version = 0;
        int runtime_version = 0;
        int l2_cache_size = 0;        // bytes
        size_t global_memory = 0;     // bytes
        size_t total_memory = 0;
        size_t free_memory = 0;
        size_t available_memory = 0;
        size_t user_requested_memory_for_allocation = 0;
        std::string name;
        std::string arch;             // e.g. "gfx950"
        int multiprocessor_count = 0;
    };

    // Selects HIP device `selected_device`; a nonzero limit caps the memory
    // the pipeline may use.  Without a usable device the object is not
    // initialised and free_memory() returns 0.
    explicit aa_device_info(int selected_device, size_t maximum_allowed_memory_for_pipeline = 0);

    // No reference counterpart: a device description that only budgets
    // memory (no HIP calls), for computing strategies on a host without a GPU.
    static aa_device_info memory_only(size_t available_bytes, int card_number = 0);

    size_t free_memory();
    bool request_memory(size_t memory_size_in_bytes);
    bool requested_memory() { return m_requested != 0; }
    bool reset_requested_memory() {
        m_requested = 0;
        return true;
    }
    int currently_selected_card_id() const { return m_device; }
    bool print_card_info();
    bool is_initialised() const { return m_init; }
    const aa_card_info &card_info() const { return m_card; }

private:
    aa_device_info() = default;
    int m_device = 0;
    bool m_init = false;
    bool m_memory_only = false;
    size_t m_limit = 0;
    size_t m_requested = 0;
    aa_card_info m_card;
};

}  // namespace astroaccelerate
#endif
