// SourceStage<Loader> / SinkStage<Saver>: the ends of a pipeline
// (glados/pipeline/SourceStage.h, SinkStage.h keep the same roles).  The
// loader's loadImage() returns frames until an invalid Image ends the
// stream, which is passed on; the saver's saveImage(img, path) consumes
// frames until the sentinel arrives.
#pragma once

#include <cstdint>
#include <string>
#include <utility>

#include "../Image.h"
#include "Port.h"

namespace glados {
namespace pipeline {

template <class Loader>
class SourceStage : public Loader, public OutputSide<Image<typename Loader::manager_type>> {
public:
    using output_type = Image<typename Loader::manager_type>;

    template <typename... Args>
    explicit SourceStage(Args&&... args) : Loader(std::forward<Args>(args)...)
    {
    }

    void run()
    {
        for (;;) {
            output_type img = Loader::loadImage();
            if (!img.valid()) {
                this->output(output_type());
                break;
            }
            this->output(std::move(img));
            ++num_;
        }
    }

    std::uint64_t num() const noexcept { return num_; }

private:
    std::uint64_t num_ = 0;
};

template <class Saver>
class SinkStage : public Saver, public InputSide<Image<typename Saver::manager_type>> {
public:
    using input_type = Image<typename Saver::manager_type>;

    template <typename... Args>
    SinkStage(const std::string& path, const std::string& prefix, Args&&... args)
        : Saver(std::forward<Args>(args)...), path_(path), prefix_(prefix)
    {
        if (!path_.empty() && path_.back() != '/') path_.push_back('/');
    }

    void run()
    {
        for (;;) {
            input_type img = this->input_queue_.take();
            if (!img.valid()) break;
            Saver::saveImage(std::move(img), path_ + prefix_);
        }
    }

private:
    std::string path_, prefix_;
};

}  // namespace pipeline
}  // namespace glados
