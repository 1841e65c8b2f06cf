#include "projection_set.hpp"

#include <map>
#include <mutex>
#include <random>

#include "common.hpp"

namespace etn {

void ProjectionSet::setProjectionCount(uint32_t count, const std::set<uint32_t>& skipped)
{
    indices_.clear();
    for (uint32_t i = 0; i < count; ++i)
        if (!skipped.count(i)) indices_.push_back(i);
    reorder(indices_);
}

void ProjectionSetMaxAngle::reorder(std::vector<uint32_t>& indices)
{
    std::vector<uint32_t> out(indices.size());
    size_t upper = indices.size() / 2, lower = 0;
    for (size_t i = 0; i < indices.size(); ++i) out[i] = (i % 2 == 0) ? indices[upper++] : indices[lower++];
    indices.swap(out);
}

void ProjectionSetBuckets::reorder(std::vector<uint32_t>& indices)
{
    constexpr int kSegments = 4;
    const int n = static_cast<int>(indices.size());
    int cursor[kSegments];
    for (int s = 0; s < kSegments; ++s) cursor[s] = s * n / kSegments;
    std::vector<uint32_t> out(indices.size());
    for (int i = 0; i < n; ++i) out[i] = indices[cursor[i % kSegments]++];
    indices.swap(out);
}

namespace {
// boost::uniform_int<>(0, hi) on a 32-bit engine: bucket rejection sampling
// (Boost >= 1.47 generate_uniform_int; the engine range 2^32-1 exceeds the target range).
uint32_t boost_uniform_int(std::mt19937& eng, uint32_t hi)
{
    if (hi == 0) return 0;
    if (hi == 0xffffffffu) return static_cast<uint32_t>(eng());
    const uint32_t brange = 0xffffffffu;
    uint32_t bucket = brange / (hi + 1u);
    if (brange % (hi + 1u) == hi) ++bucket;
    for (;;) {
        const uint32_t r = static_cast<uint32_t>(eng()) / bucket;
        if (r <= hi) return r;
    }
}
}  // namespace

void ProjectionSetRandom::reorder(std::vector<uint32_t>& indices)
{
    std::mt19937 eng;  // default seed 5489, as boost::random::mt19937
    for (size_t i = 0; i < indices.size(); ++i) indices[i] = static_cast<uint32_t>(i);
    // std::random_shuffle(first, last, rand): for i in 1..n-1 swap with rand(i + 1)
    for (size_t i = 1; i < indices.size(); ++i) {
        const size_t j = boost_uniform_int(eng, static_cast<uint32_t>(i));
        if (i != j) std::swap(indices[i], indices[j]);
    }
}

namespace {
std::mutex& registry_mutex()
{
    static std::mutex m;
    return m;
}
std::map<std::string, ProjectionIteratorFactory::Creator>& registry()
{
    static std::map<std::string, ProjectionIteratorFactory::Creator> r;
    return r;
}
}  // namespace

std::unique_ptr<ProjectionSet> ProjectionIteratorFactory::instantiate(const std::string& id)
{
    if (id == "identity") return std::make_unique<ProjectionSetIdentity>();
    if (id == "random") return std::make_unique<ProjectionSetRandom>();
    if (id == "maxangle") return std::make_unique<ProjectionSetMaxAngle>();
    if (id == "buckets") return std::make_unique<ProjectionSetBuckets>();
    std::lock_guard<std::mutex> lock(registry_mutex());
    auto it = registry().find(id);
    if (it != registry().end()) return it->second();
    fail("Illegal ProjectionSetType!");
}

void ProjectionIteratorFactory::registerScheme(const std::string& id, Creator creator)
{
    std::lock_guard<std::mutex> lock(registry_mutex());
    registry()[id] = std::move(creator);
}

}  // namespace etn
