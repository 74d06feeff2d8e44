// Pipeline: creates stages, links them and runs each on its own thread
// (glados/pipeline/Pipeline.h:41-54 keeps create / connect / run / wait).
#pragma once

#include <memory>
#include <thread>
#include <type_traits>
#include <utility>
#include <vector>

#include "Port.h"

namespace glados {
namespace pipeline {

class Pipeline {
public:
    template <class First, class Second>
    void connect(First first, Second second)
    {
        using port_type = typename First::element_type::output_type;
        static_assert(std::is_same<port_type, typename Second::element_type::input_type>::value,
                      "connected stages must agree on the image type");
        auto port = std::unique_ptr<Port<port_type>>(new Port<port_type>);
        port->attach(second);
        first->attach(std::move(port));
    }

    template <class PipelineStage, typename... Args>
    std::shared_ptr<PipelineStage> create(Args&&... args)
    {
        return std::make_shared<PipelineStage>(std::forward<Args>(args)...);
    }

    template <class S>
    void run(S stage)
    {
        threads_.emplace_back([stage] { stage->run(); });
    }

    template <class S, class... Rest>
    void run(S stage, Rest... rest)
    {
        run(stage);
        run(rest...);
    }

    void wait()
    {
        for (auto& t : threads_) t.join();
        threads_.clear();
    }

    ~Pipeline()
    {
        for (auto& t : threads_)
            if (t.joinable()) t.join();
    }

private:
    std::vector<std::thread> threads_;
};

}  // namespace pipeline
}  // namespace glados
