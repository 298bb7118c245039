// sensitivity_data.h -- SensitivityData (include/systems/sensitivity_data.h):
// dq_i/dp_j for the selected QoIs i and every parameter j, read and written as
// sensitivities[i][j].
#pragma once

#include <vector>

#include "libmesh/parameter_vector.h"
#include "libmesh/qoi_set.h"

namespace libMesh {

class System;

class SensitivityData {
 public:
  class Row {
   public:
    Row(SensitivityData &sd, unsigned int qoi) : _sd(sd), _qoi(qoi) {}
    Number &operator[](unsigned int parameter) { return _sd.derivative(_qoi, parameter); }

   private:
    SensitivityData &_sd;
    unsigned int _qoi;
  };
  class ConstRow {
   public:
    ConstRow(const SensitivityData &sd, unsigned int qoi) : _sd(sd), _qoi(qoi) {}
    const Number &operator[](unsigned int parameter) { return _sd.derivative(_qoi, parameter); }

   private:
    const SensitivityData &_sd;
    unsigned int _qoi;
  };

  SensitivityData() = default;
  void clear() { _grad_data.clear(); }
  SensitivityData(const QoISet &qoi_indices, const System &sys, const ParameterVector &parameter_vector) {
    allocate_data(qoi_indices, sys, parameter_vector);
  }
  // sensitivity_data.C: one zeroed row of Np derivatives per selected QoI of sys
  void allocate_data(const QoISet &qoi_indices, const System &sys, const ParameterVector &parameter_vector);
  const Number &derivative(unsigned int qoi_index, unsigned int parameter_index) const {
    return _grad_data.at(qoi_index).at(parameter_index);
  }
  Number &derivative(unsigned int qoi_index, unsigned int parameter_index) {
    return _grad_data.at(qoi_index).at(parameter_index);
  }
  ConstRow operator[](unsigned int qoi) const { return ConstRow(*this, qoi); }
  Row operator[](unsigned int qoi) { return Row(*this, qoi); }

 private:
  std::vector<std::vector<Number>> _grad_data;
};

}  // namespace libMesh
