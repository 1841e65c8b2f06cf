// Projection iteration schemes: the order in which the block-iterative loop
// visits projections (ettention/libReconstruction/src/model/projectionset/).
#pragma once

#include <cstdint>
#include <functional>
#include <memory>
#include <set>
#include <string>
#include <vector>

namespace etn {

// ProjectionSet (ProjectionSet.cpp:14-25): indices 0..count-1 minus the skip
// list, then the scheme's reorder().
class ProjectionSet {
public:
    virtual ~ProjectionSet() = default;
    void setProjectionCount(uint32_t count, const std::set<uint32_t>& skipped);
    uint32_t getNumberOfProjections() const { return static_cast<uint32_t>(indices_.size()); }
    uint32_t getProjectionIndex(uint32_t n) const { return indices_.at(n); }
    const std::vector<uint32_t>& indices() const { return indices_; }

protected:
    virtual void reorder(std::vector<uint32_t>& indices) = 0;

private:
    std::vector<uint32_t> indices_;
};

// ProjectionSetIdentity.cpp:17-20
class ProjectionSetIdentity final : public ProjectionSet {
protected:
    void reorder(std::vector<uint32_t>&) override {}
};

// ProjectionSetMaxAngle.cpp:16-38: alternate between the second and first half.
class ProjectionSetMaxAngle final : public ProjectionSet {
protected:
    void reorder(std::vector<uint32_t>& indices) override;
};

// ProjectionSetBuckets.cpp:16-42: four angle segments visited round robin.
class ProjectionSetBuckets final : public ProjectionSet {
protected:
    void reorder(std::vector<uint32_t>& indices) override;
};

// ProjectionSetRandom.cpp:17-38: default-seeded mt19937, boost::uniform_int,
// std::random_shuffle over 0..n-1 (the skip list is not applied).
class ProjectionSetRandom final : public ProjectionSet {
protected:
    void reorder(std::vector<uint32_t>& indices) override;
};

// ProjectionIteratorFactory::instantiate (factory/ProjectionIteratorFactory.cpp:11-23):
// built-in ids first, then schemes registered by plugins
// (PluginManager::instantiateProjectionIterator, PluginManager.cpp:151-162).
class ProjectionIteratorFactory {
public:
    using Creator = std::function<std::unique_ptr<ProjectionSet>()>;
    static std::unique_ptr<ProjectionSet> instantiate(const std::string& id);
    static void registerScheme(const std::string& id, Creator creator);
};

}  // namespace etn
