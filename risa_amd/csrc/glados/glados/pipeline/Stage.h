// Stage<Implementation>: runs a stage class inside the pipeline
// (glados/pipeline/Stage.h:39-97 keeps the same contract).  The
// implementation provides input_type, output_type, process(input_type&&) and
// wait() -> output_type.  One thread feeds process() from the input queue,
// one forwards wait()'s results to the next stage; both stop after passing
// on the invalid sentinel Image.
#pragma once

#include <thread>
#include <utility>

#include "Port.h"

namespace glados {
namespace pipeline {

template <class Implementation>
class Stage : public InputSide<typename Implementation::input_type>,
              public OutputSide<typename Implementation::output_type>,
              public Implementation {
public:
    using input_type = typename Implementation::input_type;
    using output_type = typename Implementation::output_type;

    template <typename... Args>
    explicit Stage(Args&&... args) : Implementation(std::forward<Args>(args)...)
    {
    }

    void run()
    {
        std::thread pusher(&Stage::push, this);
        std::thread taker(&Stage::take, this);
        pusher.join();
        taker.join();
    }

private:
    void push()
    {
        for (;;) {
            input_type img = this->input_queue_.take();
            const bool last = !img.valid();
            Implementation::process(std::move(img));
            if (last) break;
        }
    }

    void take()
    {
        for (;;) {
            output_type out = Implementation::wait();
            const bool last = !out.valid();
            this->output(std::move(out));
            if (last) break;
        }
    }
};

}  // namespace pipeline
}  // namespace glados
