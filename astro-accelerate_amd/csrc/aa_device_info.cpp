// HIP-backed device selection and memory budgeting
// (reference interface: include/aa_device_info.hpp:61-130).
#include <hip/hip_runtime.h>

#include "aa_device_info.hpp"
#include "aa_log.hpp"

namespace astroaccelerate {

aa_device_info::aa_device_info(int selected_device, size_t maximum_allowed_memory_for_pipeline)
    : m_device(selected_device) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
        (void)hipGetLastError();
        LOG(log_level::notice, "aa_device_info: no HIP device available.");
        return;
    }
    if (selected_device < 0 || selected_device >= count) {
        LOG(log_level::error, "aa_device_info: device " + std::to_string(selected_device) + " does not exist.");
        return;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, selected_device) != hipSuccess) return;
    if (hipSetDevice(selected_device) != hipSuccess) return;
    m_card.compute_capability_major = prop.major;
    m_card.compute_capability_minor = prop.minor;
    m_card.clock_rate = (float)prop.clockRate;
    m_card.memory_clock_rate = (float)prop.memoryClockRate;
    m_card.memory_bus_width = (float)prop.memoryBusWidth;
    m_card.card_number = selected_device;
    m_card.l2_cache_size = prop.l2CacheSize;
    m_card.global_memory = prop.totalGlobalMem;
    m_card.name = prop.name;
    m_card.arch = prop.gcnArchName;
    m_card.multiprocessor_count = prop.multiProcessorCount;
    (void)hipDriverGetVersion(&m_card.driver_version);
    (void)hipRuntimeGetVersion(&m_card.runtime_version);
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) == hipSuccess) {
        m_card.free_memory = free_b;
        m_card.total_memory = total_b;
    }
    m_limit = maximum_allowed_memory_for_pipeline == 0 ? m_card.global_memory : maximum_allowed_memory_for_pipeline;
    m_init = true;
}

aa_device_info aa_device_info::memory_only(size_t available_bytes, int card_number) {
    aa_device_info d;
    d.m_device = card_number;
    d.m_memory_only = true;
    d.m_init = true;
    d.m_limit = available_bytes;
    d.m_card.free_memory = available_bytes;
    d.m_card.total_memory = available_bytes;
    d.m_card.global_memory = available_bytes;
    d.m_card.available_memory = available_bytes;
    d.m_card.card_number = card_number;
    d.m_card.name = "memory-only";
    return d;
}

size_t aa_device_info::free_memory() {
    if (!m_init) return 0;
    size_t free_b = m_card.free_memory;
    if (!m_memory_only) {
        size_t total_b = 0;
        if (hipSetDevice(m_device) == hipSuccess && hipMemGetInfo(&free_b, &total_b) == hipSuccess)
            m_card.free_memory = free_b;
    }
    // Same rule as the reference: the pipeline limit caps what is free, and
    // memory already requested is subtracted.
    const size_t cap = free_b >= m_limit ? m_limit : free_b;
    m_card.available_memory = cap > m_requested ? cap - m_requested : 0;
    return m_card.available_memory;
}

bool aa_device_info::request_memory(size_t bytes) {
    if (m_card.available_memory > bytes) {
        m_requested += bytes;
        m_card.available_memory -= bytes;
        return true;
    }
    return false;
}

bool aa_device_info::print_card_info() {
    LOG(log_level::dev_debug, "CARD INFORMATION:");
    LOG(log_level::dev_debug, "Name:\t\t\t" + m_card.name + " (" + m_card.arch + ")");
    LOG(log_level::dev_debug, "Compute units:\t\t" + std::to_string(m_card.multiprocessor_count));
    LOG(log_level::dev_debug, "Clock rate [kHz]:\t" + std::to_string(m_card.clock_rate));
    LOG(log_level::dev_debug, "l2_cache_size [bytes]:\t" + std::to_string(m_card.l2_cache_size));
    LOG(log_level::dev_debug, "Global memory [bytes]:\t" + std::to_string(m_card.global_memory));
    LOG(log_level::dev_debug, "Free memory [bytes]:\t" + std::to_string(m_card.free_memory));
    return true;
}

}  // namespace astroaccelerate
