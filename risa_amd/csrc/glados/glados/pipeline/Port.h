// Port, InputSide and OutputSide: the typed links between stages
// (glados/pipeline/Port.h:36-52, InputSide.h, OutputSide.h keep the same
// shapes).  A Port<T> forwards to the InputSide<T> of the next stage, so
// Pipeline::connect(a, b) compiles only when a's output_type is b's
// input_type.
#pragma once

#include <memory>
#include <stdexcept>
#include <utility>

#include "../Queue.h"

namespace glados {
namespace pipeline {

template <class T>
class InputSide {
public:
    void input(T&& in) { input_queue_.push(std::move(in)); }

protected:
    Queue<T> input_queue_;
};

template <class T>
class Port {
public:
    void forward(T&& data)
    {
        if (!next_) throw std::runtime_error("Port: not attached");
        next_->input(std::move(data));
    }
    void attach(std::shared_ptr<InputSide<T>> next) noexcept { next_ = std::move(next); }

private:
    std::shared_ptr<InputSide<T>> next_;
};

template <class T>
class OutputSide {
public:
    void output(T&& out)
    {
        if (!port_) throw std::runtime_error("OutputSide: missing port");
        port_->forward(std::move(out));
    }
    void attach(std::unique_ptr<Port<T>>&& port) noexcept { port_ = std::move(port); }

protected:
    std::unique_ptr<Port<T>> port_;
};

}  // namespace pipeline
}  // namespace glados
